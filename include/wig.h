/*
 * wig.h — C ABI of libwig.so, the MI355X (gfx950) liquid step.
 *
 * The liquid step of worldinagrain (FLIP/APIC on a MAC grid with an MG-PCG
 * pressure solve), behind the boundary the reference's host encodes it at:
 * `Simulation.encodeFluid(pass, offset)` (js/sim.js:2466-2612).  The world's
 * device buffers keep the reference's byte layouts (the liquid's WGSL
 * bindings, shaders/fluid.wgsl:108-181,2543; js/sim.js:741-786), so a
 * snapshot part uploaded here and downloaded again is the same bytes a
 * WebGPU build holds.
 *
 * Plain C types only.  Every call returns a status: 0 ok, > 0 a warning
 * (WIG_WARN_*), < 0 an error (WIG_E_*); wig_last_error() has the message
 * (kept per thread).  One world handle per host thread; calls on a handle
 * are ordered on its HIP stream and return without waiting unless stated.
 */
#ifndef WIG_H
#define WIG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WIG_OK 0
#define WIG_WARN_OVERFLOW 1       /* more particles than capacity: those past it lost */
#define WIG_NOT_READY 2           /* wig_report_take without waiting: that report has not landed yet */
#define WIG_E_INVALID (-1)        /* bad argument, size or unsupported control */
#define WIG_E_NOMEM (-2)          /* device allocation failed */
#define WIG_E_DEVICE (-3)         /* a HIP call failed (device lost, launch error) */
#define WIG_E_NOT_BUILT (-4)      /* no gfx950 code object in this library */

/* Per-step uniform, 80 bytes: `Params` (shaders/common.wgsl:5-42), as the
 * host's writeSlot fills it (js/sim.js:1972-2001).  The liquid step reads
 * size, frame, seed, dx and earthG. */
typedef struct WigParams {
    uint32_t size[2];
    uint32_t frame;
    uint32_t seed;
    uint32_t offset[2];
    float supportMax;
    float dx;
    float day;
    float dayRate;
    float holePull;
    float massPull;
    float earthG;
    float dt;
    float dayDt;
    float clockSpan;
    uint32_t heatStages;
    float heatLimit;
    uint32_t holding;
    float pad;
} WigParams;

/* Stop criteria of the pressure solve. */
#define WIG_TOL_ABS_MAX 0   /* max |r| < tol, the reference's LIQUID_TOLERANCE (fluid.wgsl:2221-2226) */
#define WIG_TOL_REL_L2 1    /* ||r||_2 <= tol * ||b||_2 */

/* What the host decides per step (js/sim.js:2466-2612), plus the solver
 * settings the reference fixes as constants (js/sim.js:61-63,
 * fluid.wgsl:2063,2392). */
typedef struct WigStepCtl {
    uint32_t full;           /* fluidStats.active || bodiesRun(); 0 takes the idle path (:2478-2485) */
    uint32_t viscous;        /* fluidStats.viscous || surfaceTension(): the viscosity's sweeps run
                                (:3277-3466; js/sim.js:2505-2520) */
    uint32_t surface;        /* surfaceTension(): the surface pass runs (:1698-1807; js/sim.js:2449-2497) */
    uint32_t bodies_run;     /* bodiesRun(): what the liquid did to the bodies is handed to them (rigidify,
                                bodyPush: :2944-3218, js/sim.js:2594-2599); needs `bodies` */
    uint32_t fluid_cycles;   /* PCG iterations encoded (sim.js fluidCycles, <= WIG_MAX_CYCLES) */
    uint32_t shifted_first;  /* tiling of the step's first relaxation (sim.js shiftedFirst after its toggle) */
    uint32_t tol_mode;       /* WIG_TOL_* */
    float tol;               /* 1e-2 with WIG_TOL_ABS_MAX reproduces the reference */
    uint32_t pre_relax;      /* relaxation dispatches before the coarse levels (PRE_RELAX = 1) */
    uint32_t post_relax;     /* ... and after (POST_RELAX = 1) */
    uint32_t sweep_pairs;    /* red-black sweep pairs per relaxation dispatch (SWEEPS = 4) */
    uint32_t coarse_sweeps;  /* SOR sweep pairs on the coarsest level (COARSE_SWEEPS = 40) */
    uint32_t first_chunk;    /* 0: encode fluid_cycles iterations, the reference's way.  Else this
                                library's: encode this many, then wait for the solve's convergence flag
                                and encode chunk more at a time until it converges or fluid_cycles run */
    uint32_t chunk;
    uint32_t level_pairs;    /* red-black sweep pairs on each coarse level, down and up (0: 1, the reference's) */
    uint32_t vcycle_storage; /* WIG_STORE_*: what the V-cycle (the preconditioner) keeps its values in */
    uint32_t residual_every; /* 0: the reference's recursive residual throughout.  k > 0 (this library's,
                                at most WIG_MAX_CYCLES): after every k-th iteration the residual is
                                recomputed as b - A x (FP64 accumulation) and replaces the recursive one,
                                and the stop is judged only there, so a solve stops on its true residual */
    uint32_t air;            /* 1: the air couples in, as the reference's bindings air[0] and airSource
                                (fluid.wgsl:116-181): WIG_PART_AIR's pressure pushes surface particles
                                (gather, :1027-1032), its wind drags the surface faces (setFaces,
                                :1874-1932), and what the step gives the air goes into
                                WIG_PART_AIR_SOURCE (the drag's pushes; the gas the liquid displaces,
                                raster's pushGasOut, :3855-3861).  0: the air at rest at uniform
                                pressure, nothing added to WIG_PART_AIR_SOURCE (displaced gas is let go) */
    float surface_drag;      /* with air: the wind's drag coefficient along the surface,
                                (KARMAN / ln(max(dx / 2 / 1e-4, 2)))^2 (fluid.wgsl:1884-1895); the host
                                works it out from dx (worldinagrain_amd/host.py) so every backend takes
                                the same float */
    uint32_t occupants;      /* 1: people's bodies, WIG_PART_OCCUPANTS, are obstacles that move
                                (fluid.wgsl:175-178): a particle's or the density splat's wall
                                (:424-431, :1056-1062), an obstacle cell to classify (:1123-1125), the
                                faces beside one move at the person's velocity (obstacleVel,
                                :1820-1827), no gas is claimed into one (:3919, :3937), and raster leaves
                                its particles undrawn and its liquid turned air (:3995-4004).  0: no
                                one (the part is not read) */
    uint32_t gravity;        /* 1: the gravity pass's field, WIG_PART_GRAVITY, gives the pull and down
                                (pull.wgsl:13-44, read through fluid.wgsl's binding 10): the faces'
                                pull (setFaces, :1968-1972), the slip and the falling test (advect,
                                :3613,3686), and down for the bubble test, the escape and where
                                displaced gas goes (:1138, :3577, :3822-3825, :3912).  0: Earth's pull
                                alone, (0, earthG), down (0, 1) (the part is not read) */
    uint32_t bodies;         /* 1: free and held bodies, as the reference's bindings 12, 14 and 16
                                (fluid.wgsl:1181-1475): WIG_PART_MOTION's table and WIG_PART_SWEPT's
                                words are read, a loose structural solid (its element's `breaks`) of an
                                alive and weighed body is the body's cell to classify, its faces move
                                with the body, the solve takes its give (A + J^T M^-1 J: cgResidual,
                                cgApply), the projection moves its faces with the push found, and
                                WIG_PART_PUSHES holds the step's push sums.  0: no body anywhere (the
                                parts are not read).  Not with residual_every (WIG_E_INVALID) */
} WigStepCtl;

/* V-cycle storage: float in the reference's buffers (results bit-identical
 * to the reference's arithmetic), or 16 bits a value (IEEE half or
 * bfloat16, arithmetic in float) in buffers of this library's own; the
 * outer PCG stays float with FP64 sums either way. */
#define WIG_STORE_F32 0
#define WIG_STORE_F16 1
#define WIG_STORE_BF16 2

/* first_chunk: as many iterations as the world's last solve ran */
#define WIG_CHUNK_AUTO 0xffffffffu

#define WIG_MAX_CYCLES 256

/* The world's buffers, by the reference's part names (js/passes.js:78-84). */
#define WIG_PART_CELLS 0       /* cells[cur]: 16 B a cell, then the gas plane (8 B a cell) */
#define WIG_PART_PARTICLES0 1  /* `sorted`: 16 B header, then 3 vec4 streams of cap particles */
#define WIG_PART_PARTICLES1 2  /* `parts`, the step's */
#define WIG_PART_COUNTERS 3    /* 11 u32 a cell, then 2 a face */
#define WIG_PART_FIELD 4       /* vec4<f32>: 2 a cell, 1 a face, viscosity words, MG levels */
#define WIG_PART_BLOCKS 5      /* per block marks, state, lists, FLUID_STATS, totals */
#define WIG_PART_FLUIDCG 6     /* vec4<f32> per cell (x, r, p, Ap), per block, sums */
#define WIG_PART_SOLVE 7       /* this library's FP64 solve state and solver words (no counterpart in the reference) */
#define WIG_PART_AIR 8         /* air[0]: a vec4<f32> an air cell (AIR_CELL 1: a world cell): gauge pressure
                                  (P_ATM 12 an atmosphere), the velocity on its right face and on its bottom
                                  face (cells a step), temperature or AIR_BLOCKED (-1e4: no gas) or AIR_SOLID
                                  (-2e4) (common.wgsl:280-288; air.wgsl:1742-1767); read */
#define WIG_PART_AIR_SOURCE 9  /* airSource: 16 u32 an air cell, exact sums of two words (low, high) in
                                  units of 2^-32: the velocity added to its right and bottom faces, then
                                  gas amount, vapour, mist, oxygen lost, heat (airsource.wgsl:10-33);
                                  added to by the step, taken in and cleared by the air's own passes */
#define WIG_PART_OCCUPANTS 10  /* occupants: a u32 a cell, common.wgsl's occupantWord (:238-254): 0 nobody,
                                  else the person's slot + 1 in the low byte and their velocity times 128
                                  (cells a step) in two 12-bit two's-complement fields, x at bit 8 and y at
                                  bit 20; read */
#define WIG_PART_GRAVITY 11    /* the gravity pass's field (a texture_2d<f32> in the reference): a vec4<f32>
                                  a light cell of 2 x 2 cells (GRAVITY_BLOCK), ceil(W/2) x ceil(H/2) of them,
                                  row-major: the masses' pull (x, y; cells a step squared), their potential
                                  (z), and down (w: the step index, 0..7, plus 8 where the pull is too weak
                                  to settle anything) (pull.wgsl:1-44); read */
#define WIG_PART_MOTION 12     /* bodyMotion (binding 12): the `motion` buffer's words up to the end of its
                                  liquid slots, WIG_MOTION_WORDS u32 (bodies.wgsl's layout: MAX_BODIES
                                  words, then the bodies' table, MAX_BODIES slots of 116 words (js/shared.js's
                                  BODY_WORD: pose, velocity, spin, mass, inertia, cells, flags, subcell,
                                  cargo, mobility), then the contacts, then 4 words a body of what the
                                  liquid did to it); the table read, the liquid slots added to
                                  (fluid.wgsl:1211-1228) */
#define WIG_PART_SWEPT 13      /* the bodies buffer's swept words (binding 14 from common.wgsl's
                                  bodySweptSlot(size, 0, 0) on): two u32 a cell, for even and odd steps, the
                                  body that last moved over it (common.wgsl:355-370); read */
#define WIG_PART_PUSHES 14     /* bodyPushes (binding 16): WIG_BODY_PUSH_WORDS u32 a body, six sets of exact
                                  sums of a force and a torque (fluid.wgsl:1304-1340); the step's own scratch,
                                  emptied as it starts */
#define WIG_PART_COUNT 15

/* The bodies (js/shared.js:299-333). */
#define WIG_MAX_BODIES 1024u
#define WIG_BODY_PUSH_WORDS 36u
#define WIG_MOTION_WORDS (WIG_MAX_BODIES * (1u + 116u) + 16384u * 12u + WIG_MAX_BODIES * 4u)

/* FLUID_STATS words (js/shared.js:271-272). */
#define WIG_STAT_ACTIVE 0
#define WIG_STAT_RETIRED 1
#define WIG_STAT_VISCOUS 2
#define WIG_STAT_PARTICLES 3
#define WIG_STAT_RESIDUAL 4
#define WIG_STAT_CYCLES 7
#define WIG_STAT_MOST_CYCLES 8
#define WIG_STATS 9

/* What the last step's solve did, from WIG_PART_SOLVE. */
typedef struct WigSolveStats {
    double rz, alpha, beta, pq;   /* the last iteration's r.z, alpha, beta, p.Ap */
    double rr;                    /* ||r||^2 after the last iteration run */
    double bb;                    /* ||b||^2 of the step's right-hand side */
    uint32_t cycles;              /* iterations run */
    uint32_t converged;
    float vcycle_scale;           /* 16-bit V-cycle storage: the last power of two its rhs was scaled by */
    uint32_t pad;
} WigSolveStats;

/* A world of width x height cells (the reference's Simulation buffers,
 * js/sim.js:741-786).  coarse_levels 0 takes the reference's rule (halve
 * until <= 512 cells, at most 8: js/sim.js:763-770); otherwise that many
 * coarse levels.  Replaces `new Simulation(device, pipes, w, h, ...)`'s
 * liquid allocation (js/sim.js:664,741-786).  NULL on error. */
typedef struct WigWorld WigWorld;
WigWorld *wig_create(uint32_t width, uint32_t height, uint32_t coarse_levels, uint32_t flags);

/* Frees the world's device buffers (js/sim.js:1502-1527, Simulation.destroy). */
int wig_destroy(WigWorld *w);

/* The element kinds the liquid step reads (elems[t].kind, js/elements.js:8:
 * 1 solid, 2 powder, 3 liquid, 4 gas), for element ids 0..255
 * (the `elems` binding, fluid.wgsl:117). */
int wig_set_kinds(WigWorld *w, const uint8_t kinds[256]);

/* The elements' columns the liquid step reads beside their kind
 * (common.wgsl's Elem; js/elements.js's table): a liquid's viscosity
 * (cells^2 a step) at viscosity_t (degrees C) and its temperature scale
 * viscosity_e (degrees C; <= 0: the same at every temperature), as
 * viscosityAt (common.wgsl:1486-1493) takes them; the surface tension
 * factor (surface, fluid.wgsl:1779); a solid's wetting share (:1781); and
 * whether a structural solid can come loose as a free body's.  All zero
 * until set: no viscosity, no tension, no loose solid. */
typedef struct {
    float viscosity, viscosity_e, viscosity_t, tension, wetting;
    float breaks;   /* 1: elems[t].breakTo != NO_ELEM, a structural solid that can come loose
                       in a free body (fluid.wgsl:1233-1236); 0: not */
} WigElementProps;
int wig_set_element_props(WigWorld *w, const WigElementProps props[256]);

/* The world's sizes: cells, faces, blocks, particle capacity, coarse levels. */
typedef struct WigLayoutInfo {
    uint32_t width, height, cells, faces, blocks, blocks_wide, capacity, levels;
    uint32_t level_width[9], level_height[9];
    uint64_t part_bytes[WIG_PART_COUNT];
} WigLayoutInfo;
int wig_layout(const WigWorld *w, WigLayoutInfo *out);

/* Copies a part to or from the device (js/sim.js:1429-1493's snapshot
 * parts; queue.writeBuffer / mapAsync).  Synchronous.  The parts of the
 * step's optional inputs (air, airSource, occupants, gravity and the
 * bodies') take device memory only once uploaded, asked for by
 * wig_device_ptr, or stepped with their WigStepCtl switch on; a download
 * before that reads zeros. */
int wig_upload(WigWorld *w, int part, const void *src, size_t bytes);
int wig_download(WigWorld *w, int part, void *dst, size_t bytes);

/* The device address of a part, for callers that share memory with the
 * world (bytes of it as wig_layout gives them). */
int wig_device_ptr(WigWorld *w, int part, void **ptr);

/* Runs later calls on this HIP stream (a hipStream_t; NULL: the world's own).
 * They are ordered after the work already enqueued on the stream the world
 * ran on (the new stream waits for an event recorded on the old). */
int wig_set_stream(WigWorld *w, void *stream);

/* One liquid step, enqueued (Simulation.encodeFluid, js/sim.js:2466-2612). */
int wig_fluid_step(WigWorld *w, const WigParams *params, const WigStepCtl *ctl);

/* FLUID_STATS as the frame's report copies them (js/sim.js:1760-1780), then
 * MOST_CYCLES cleared as a report on steps does.  Waits for the stream. */
int wig_report(WigWorld *w, uint32_t stats[WIG_STATS]);

/* The same report asked for without waiting, as the reference maps its
 * report buffer asynchronously and takes it in when it lands
 * (Simulation.readReport's mapAsync, js/sim.js:1882-1901): the words are
 * copied behind the work already enqueued into one of two pinned host
 * slots, and *ticket names them.  wig_report_take copies them out: with
 * `wait` once they have landed, else at once or WIG_NOT_READY.  The two
 * latest tickets are valid (a third reuses the first's slot); older ones,
 * and a batch's tickets from before wig_batch_assign, are WIG_E_INVALID. */
int wig_report_async(WigWorld *w, uint64_t *ticket);
int wig_report_take(WigWorld *w, uint64_t ticket, uint32_t stats[WIG_STATS], int wait);

/* The last solve's FP64 scalars.  Waits for the stream. */
int wig_solve_stats(WigWorld *w, WigSolveStats *out);

/* Running totals over every full step since the world was made (this
 * library's own counters, no counterpart in the reference): the steps, the
 * PCG iterations run, the particles stepped and the cells of the blocks
 * worked on, summed on the device so a timed loop reads them only after. */
typedef struct WigRunTotals {
    uint64_t steps, cycles, particles, active_cells;
} WigRunTotals;
int wig_run_totals(WigWorld *w, WigRunTotals *out);

/* Kernel timing (this library's, for measurement): with id >= 0, every later
 * launch of that kernel is bracketed by HIP events on the world's stream;
 * wig_kernel_time waits and gives the launches and their summed time, and
 * clears them.  id < 0 stops.  Ids are WIG_K_*; wig_kernel_name names one. */
int wig_time_kernel(WigWorld *w, int id);
int wig_kernel_time(WigWorld *w, double *ms, uint32_t *launches);
const char *wig_kernel_name(int id);
int wig_kernel_count(void);

/* The tile side (8, 16 or 32 cells) of the coarse-level sweep kernels on
 * coarse level `level`, for the levels swept by tiles rather than in one
 * workgroup (1 .. the level where that begins; WIG_E_INVALID on others).
 * Results do not depend on it, only time does: the autotuner
 * (worldinagrain_amd/tune.py) sets it.  This library's own; the reference
 * sweeps a whole level with 8 x 8 workgroups (fluid.wgsl:2317-2387). */
int wig_set_tile(WigWorld *w, uint32_t level, uint32_t side);

/* The coarse level from which one workgroup solves the rest of the V-cycle
 * in LDS (the levels above it are swept in tiles): from the first level
 * LDS can hold with all below it (the default) down to the coarsest.
 * Results do not depend on it either; the autotuner sets it.  This
 * library's own. */
int wig_set_lds_level(WigWorld *w, uint32_t level);

/* Within that workgroup, the coarse levels of at most `cells` cells below
 * its first are worked by one wavefront, without the workgroup's barriers
 * (0: none; 256 by default).  Results do not depend on it; the autotuner
 * sets it.  This library's own. */
int wig_set_coarse_wave(WigWorld *w, uint32_t cells);

/* Waits for everything enqueued on the world's stream. */
int wig_sync(WigWorld *w);

/* The FP64 check of a pressure solve (this library's own; the reference has
 * none).  With capture on, each solve keeps its right-hand side b.
 * wig_cg64 then solves A x64 = b by plain CG in FP64 on the device, on the
 * last solve's own matrix (its face betas and pins, as the solve applies
 * them), from 0 to ||b - A x64|| <= tol ||b|| or maxit iterations, and
 * measures the solve's float pressure x against it: err = ||x - x64|| /
 * ||x64|| and relres = ||b - A x|| / ||b||, both in FP64 over the fluid
 * cells.  Waits. */
typedef struct WigCheck64 {
    uint32_t iterations, pad;
    double relres64;    /* of x64 */
    double err;         /* of the solve's x against x64 */
    double relres;      /* of the solve's x */
} WigCheck64;
int wig_capture(WigWorld *w, int on);
int wig_cg64(WigWorld *w, double tol, uint32_t maxit, WigCheck64 *out);
/* x64, one double a cell (0 off the fluid), after wig_cg64. */
int wig_cg64_x(WigWorld *w, double *x);

/* Batches: many worlds, of any sizes, stepped together, every kernel of the
 * step one launch over all of them (this library's own; the reference steps
 * one Simulation a frame).  The worlds stay the caller's: each must outlive
 * its use by the batch, until the batch is destroyed or assigned other
 * worlds.  A world's results do not depend on the batch it is stepped in
 * (each runs the solver iterations its own host would encode).  The batch
 * runs on a stream of its own unless set otherwise (wig_batch_set_stream;
 * NULL: its own again); a world stepped both alone and in batches, or
 * copied on its own stream, is the caller's to order. */
typedef struct WigBatch WigBatch;
WigBatch *wig_batch_create(WigWorld *const *worlds, uint32_t n);
int wig_batch_destroy(WigBatch *b);
/* The batch's worlds replaced by these n (a server's next round of worlds). */
int wig_batch_assign(WigBatch *b, WigWorld *const *worlds, uint32_t n);
int wig_batch_set_stream(WigBatch *b, void *stream);
/* One liquid step of each world i, with params[i] and ctl[i] (as wig_fluid_step). */
int wig_batch_step(WigBatch *b, const WigParams *params, const WigStepCtl *ctl);
/* stats[i * WIG_STATS ...]: each world's report (as wig_report).  Waits. */
int wig_batch_report(WigBatch *b, uint32_t *stats);
/* The batch's reports without waiting, as wig_report_async / wig_report_take. */
int wig_batch_report_async(WigBatch *b, uint64_t *ticket);
int wig_batch_report_take(WigBatch *b, uint64_t ticket, uint32_t *stats, int wait);
int wig_batch_sync(WigBatch *b);
/* hipGraphs: with `on`, each phase of a step (the passes before the solve,
 * its start, one iteration, its end, the rest) is captured once into a
 * graph on its first run and replayed after, until the launch layout changes
 * (world tiles, LDS sizes, the stream).  The stream must be one of the
 * caller's or the world's own, not the null stream.  Off while a kernel is
 * timed.  This library's own. */
int wig_set_graphs(WigWorld *w, int on);
int wig_batch_set_graphs(WigBatch *b, int on);
/* Kernel timing over the batch's launches, as wig_time_kernel / wig_kernel_time. */
int wig_batch_time_kernel(WigBatch *b, int id);
int wig_batch_kernel_time(WigBatch *b, double *ms, uint32_t *launches);

/* ------------------------------------------------------------------
 * The air's projection: the MG-PCG that solves for the air's dynamic
 * pressure each step before the liquid step (shaders/air.wgsl:869-1536),
 * on the reference's buffers.  An air grid of width x height air cells
 * (AIR_CELL 1: the world's cells) with its levels: halved, rounding up,
 * until a level holds 256 cells or fewer, 9 levels at most
 * (js/sim.js:712-719); a grid whose 9th level would hold more is refused.
 * Replaces the projection's share of `new Simulation(...)`'s air
 * allocation (js/sim.js:705-723,858-864). */
typedef struct WigAir WigAir;
/* Parts (the reference's buffers):
 *   MG      airProjection: every level, two float4s a cell, level l at
 *           level_base[l] float4s (air.wgsl:84-90,172-192);
 *   CG      airCg: a float4 a level-0 cell (x, r, p, A p), then the sums,
 *           four to a float4: r.z, alpha, beta, a spare, a partial per
 *           8 x 8 group (air.wgsl:947-961);
 *   STATUS  airFastest's 8 words (air.wgsl:96-104): [1..3] the
 *           iterations' most |r| in turn, [4] the iterations run;
 *   MEAN    a float per level-0 cell: meanExpansion of its zone
 *           (air.wgsl:1006-1021), which cgStart adds to the divergence of
 *           cells with gas; the zones are the caller's (none: zeros). */
#define WIG_AIR_PART_MG 0
#define WIG_AIR_PART_CG 1
#define WIG_AIR_PART_STATUS 2
#define WIG_AIR_PART_MEAN 3
#define WIG_AIR_PART_COUNT 4
typedef struct WigAirLayout {
    uint32_t width, height, levels, groups;
    uint32_t level_width[9], level_height[9], level_base[9];
    uint64_t part_bytes[WIG_AIR_PART_COUNT];
} WigAirLayout;
/* Sides 1 .. 16384; NULL on error (wig_last_error has why). */
WigAir *wig_air_create(uint32_t width, uint32_t height);
int wig_air_destroy(WigAir *a);
int wig_air_layout(const WigAir *a, WigAirLayout *out);
/* Copies a part to or from the device, whole.  Synchronous. */
int wig_air_upload(WigAir *a, int part, const void *src, size_t bytes);
int wig_air_download(WigAir *a, int part, void *dst, size_t bytes);
int wig_air_device_ptr(WigAir *a, int part, void **ptr);
/* Runs later calls on this HIP stream (NULL: the grid's own), ordered after
 * the work already enqueued on the old one, as wig_set_stream. */
int wig_air_set_stream(WigAir *a, void *stream);
/* The side of a level's LDS tiles in the V-cycle's passes, 8, 16 or 32
 * (default 32 on level 0, 16 on level 1, 8 below): the results are the same bits
 * either way.  This library's own. */
int wig_air_set_tile(WigAir *a, uint32_t level, uint32_t side);
/* With `on`, the levels from the first that fits 64 x 64 cells down to the
 * coarsest and back run in one workgroup's kernel; off (the default), a
 * launch each.  The same bits either way.  This library's own. */
int wig_air_set_coarse(WigAir *a, int on);
/* The projection's solve, enqueued as encodeAir dispatches it
 * (js/sim.js:2264-2289): gather's clearing of the solve's status slots
 * (air.wgsl:444-449), cgStart, then `cycles` iterations (the host's
 * airCycles, js/sim.js:36,1919), each a V-cycle, cgRho, cgBeta,
 * cgDirection, cgApply, cgAlpha, cgUpdate; an iteration after the most |r|
 * left falls under AIR_TOLERANCE (1e-4) changes nothing but the slots.
 * Level 0 of MG (gather's and divergence's, air.wgsl:352-414,825-866) and
 * CG's potentials (last step's) are the input.  `cycles` at most 4096
 * (WIG_E_INVALID above). */
int wig_air_project(WigAir *a, uint32_t cycles);
int wig_air_sync(WigAir *a);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc): wig_air_project's projection, the same bits, on buffers the
 * caller holds, laid out as wig_air_layout gives them for a grid of
 * width x height (mg, cg, status, mean: the four parts), enqueued on
 * `stream` (a hipStream_t; NULL: the null stream).  It allocates nothing:
 * its working arrays (4 KB + 32 B a cell) live in a static device array of
 * the library's, one per device, held until the process ends, sized for
 * WIG_AIR_STATIC_MAX_CELLS cells (WIG_E_INVALID above).  Every call on a
 * device shares that device's array, so calls that overlap on one device
 * (two streams, or two threads) race: order them.  The tile sides are the
 * defaults and the coarse levels run a launch each. */
#define WIG_AIR_STATIC_MAX_CELLS 262144u
int wig_air_project_static(void *mg, void *cg, void *status, const void *mean, uint32_t width, uint32_t height,
                           uint32_t cycles, void *stream);
/* With `on`, each wig_air_project is timed with HIP events on its stream
 * (and waits for its end); wig_air_elapsed has the sum and the count. */
int wig_air_time(WigAir *a, int on);
int wig_air_elapsed(WigAir *a, double *ms, uint32_t *calls);

/* The last error on this thread, "" if none. */
const char *wig_last_error(void);

/* 1 if the library holds gfx950 code and a device is visible. */
int wig_device_ok(void);

#ifdef __cplusplus
}
#endif
#endif
