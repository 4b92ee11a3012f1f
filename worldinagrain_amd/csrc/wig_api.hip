// The C ABI of libwig.so (include/wig.h): world buffers in the reference's
// layouts on the device, and one liquid step enqueued in encodeFluid's order
// (js/sim.js:2466-2612).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/wig.h"
#include "wig_device.h"

// the kernels (wig_kernels.hip)
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v);
__global__ void k_sync(WorldGrid g);
__global__ void k_syncParticles(WorldGrid g);
__global__ void k_syncSeed(WorldGrid g);
__global__ void k_compactMarks(WorldGrid g);
__global__ void k_compact(WorldGrid g);
__global__ void k_retire(WorldGrid g);
__global__ void k_clear(WorldGrid g);
__global__ void k_gather(WorldGrid g);
__global__ void k_localJoin(WorldGrid g, int mode);
__global__ void k_edgeJoin(WorldGrid g, int mode);
__global__ void k_sealAnchor(WorldGrid g);
__global__ void k_classify(WorldGrid g);
__global__ void k_setFaces(WorldGrid g);
__global__ void k_viscosity(WorldGrid g, int color);
__global__ void k_surface(WorldGrid g);
__global__ void k_divergence(WorldGrid g);
template <int S> __global__ void k_relax(WorldGrid g, int post);
template <int S> __global__ void k_restrict(WorldGrid g);
template <int S> __global__ void k_levelRho(WorldGrid g);
template <int S> __global__ void k_levelDown(WorldGrid g, uint32_t l);
template <int S> __global__ void k_coarseLDS(WorldGrid g);
template <int S> __global__ void k_levelUp(WorldGrid g, uint32_t l);
__global__ void k_cgStart(WorldGrid g);
__global__ void k_cgAnchor(WorldGrid g);
__global__ void k_cgResidual(WorldGrid g);
__global__ void k_sumNorm(WorldGrid g);
template <int S> __global__ void k_cgPrepare(WorldGrid g);
template <int S> __global__ void k_cgRho(WorldGrid g);
__global__ void k_cgBeta(WorldGrid g);
template <int S> __global__ void k_cgApply(WorldGrid g);
__global__ void k_cgAlpha(WorldGrid g);
template <int S> __global__ void k_cgUpdate(WorldGrid g);
template <int S> __global__ void k_cgTrue(WorldGrid g);
__global__ void k_cgCheck(WorldGrid g);
template <int S> __global__ void k_cgFinish(WorldGrid g);
__global__ void k_bodyStart(WorldGrid g);
__global__ void k_bodyLoads(WorldGrid g);
__global__ void k_bodyDrag(WorldGrid g);
__global__ void k_bodyFaces(WorldGrid g);
__global__ void k_cgBodyX(WorldGrid g);
template <int S> __global__ void k_cgBodyP(WorldGrid g);
__global__ void k_cgBodyF(WorldGrid g);
__global__ void k_rigidify(WorldGrid g);
__global__ void k_bodyPush(WorldGrid g);
__global__ void k_project(WorldGrid g);
__global__ void k_extrapolate(WorldGrid g, int layer);
__global__ void k_advect(WorldGrid g);
__global__ void k_sortCount(WorldGrid g);
__global__ void k_scanBlocks(WorldGrid g);
__global__ void k_scanTotals(WorldGrid g);
__global__ void k_scanFinish(WorldGrid g);
__global__ void k_scatter(WorldGrid g);
__global__ void k_rasterClaims(WorldGrid g);
__global__ void k_raster(WorldGrid g);
__global__ void k_unpaint(WorldGrid g);
__global__ void k_stops(WorldGrid g, uint32_t *out);
__global__ void k64_start(WorldGrid g);
__global__ void k64_sum(WorldGrid g, int which);
__global__ void k64_apply(WorldGrid g);
__global__ void k64_update(WorldGrid g);
__global__ void k64_direction(WorldGrid g);
__global__ void k64_check(WorldGrid g, int phase);
__global__ void k64_residual(WorldGrid g, int of_x64);
__global__ void k64_keep(WorldGrid g, int nv, int at);
__global__ void k_reports(WorldGrid g, uint32_t *out);

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
// (the air's projection's entry points, wig_air.hip, report through this)
int wig_fail_message(int code, const char *msg) { return fail(code, "%s", msg); }
#define HIPCHECK(x)                                                                         \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return fail(WIG_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// Kernel ids for wig_time_kernel, in launch order.
#define WIG_KERNELS(X) X(sync) X(syncParticles) X(syncSeed) X(compactMarks) X(compact) X(retire) X(clear) X(gather) \
    X(localJoin) X(edgeJoin) X(sealAnchor) X(classify) X(surface) X(bodyStart) X(bodyLoads) X(setFaces) X(viscosity) \
    X(bodyDrag) X(bodyFaces) X(divergence) X(cgStart) X(cgAnchor) X(cgBodyX) \
    X(cgResidual) X(sumNorm) X(cgPrepare) X(levelRho) X(relax) X(restrict) X(levelDown) X(coarseLDS) X(levelUp) \
    X(cgRho) X(cgBeta) X(cgBodyP) X(cgApply) X(cgAlpha) X(cgUpdate) X(cgTrue) X(cgCheck) \
    X(cgFinish) X(cgBodyF) X(project) X(rigidify) X(bodyPush) X(extrapolate) X(advect) X(sortCount) X(scanBlocks) \
    X(scanTotals) X(scanFinish) X(scatter) X(rasterClaims) X(raster) X(unpaint)
#define X_ENUM(n) KID_##n,
enum { WIG_KERNELS(X_ENUM) KID_COUNT };
#define X_NAME(n) #n,
static const char *g_kernel_names[] = {WIG_KERNELS(X_NAME)};
#define MAX_TIMED 8192

// Launch shapes: how many workgroups a kernel takes for one world.
enum {
    SH_ONE,      // one (the single-workgroup passes)
    SH_BG,       // 64-thread passes striding over the active blocks
    SH_BG4,      // 256-thread passes, four blocks a workgroup
    SH_PG,       // 256-thread passes over the particles
    SH_WORLD,    // one workgroup a block, bw x bh
    SH_LVL1,     // 8 x 8 cells of coarse level 1 a workgroup
    SH_B256,     // 256 blocks a workgroup
    SH_BG16,     // 256-thread passes, sixteen blocks a workgroup (four a wavefront)
    SH_RELAX,    // k_relax: 4 x RELAX_NT tiles a 256-thread workgroup
    SH_TILE,     // + l - 1: k_levelDown/k_levelUp's tiles on level l
    SH_COUNT = SH_TILE + MG_LEVELS
};

static uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }

// k_levelDown/k_levelUp: tiles of T x T level cells, each with a halo of
// one cell a half-sweep (k_levelDown one more), in LDS 7 floats a cell, and
// k_levelDown's residuals (2 a tile cell) or k_levelUp's coarse cells (3)
#define MAX_LEVEL_PAIRS 8u
static size_t downLds(uint32_t pairs, uint32_t T) {
    size_t R = T + 2u * (2u * pairs + 1u);
    return 4u * (7u * R * R + 2u * T * T);
}
static size_t upLds(uint32_t pairs, uint32_t T) {
    size_t R = T + 4u * pairs, Rc = R / 2u + 3u;
    return 4u * (7u * R * R + 3u * Rc * Rc);
}
// a tile side for level_pairs sweep pairs, narrowed to fit 64 KB of LDS
static uint32_t fitTile(uint32_t T, uint32_t pairs) {
    while (T > 8u && std::max(downLds(pairs, T), upLds(pairs, T)) > 65536u) T /= 2u;
    return T;
}

// Grid of the passes over the blocks worked on: they stride over the active
// list, so a few workgroups a CU cover any count.
static uint32_t block_grid(const Dev &d) { return std::min<uint32_t>(d.B + d.bw + d.bh + 1u, 4096u); }
static uint32_t particle_grid(const Dev &d) { return std::min<uint32_t>((d.cap + 255u) / 256u, 4096u); }

static uint32_t shape_count(const Dev &d, int sh) {
    switch (sh) {
    case SH_ONE: return 1u;
    case SH_BG: return block_grid(d);
    case SH_BG4: return std::min<uint32_t>((d.B + 3u) / 4u, 1024u);
    case SH_PG: return particle_grid(d);
    case SH_WORLD: return d.bw * d.bh;
    case SH_LVL1: return ceil_div(d.lw[1], 8u) * ceil_div(d.lh[1], 8u);
    case SH_B256: return ceil_div(d.B, 256u);
    case SH_BG16: return std::min<uint32_t>((d.B + 15u) / 16u, 1024u);
    case SH_RELAX: return std::min<uint32_t>((d.B + d.bw + d.bh + 4u * RELAX_NT) / (4u * RELAX_NT), 1024u);
    default: {
        uint32_t l = (uint32_t)(sh - SH_TILE) + 1u;
        if (l >= d.lds_level) return 0u;
        return ceil_div(d.lw[l], d.tile[l]) * ceil_div(d.lh[l], d.tile[l]);
    }
    }
}

struct Batch;

struct WigWorld {
    Dev d;                       // sizes and device pointers (the step's words are filled per step)
    uint64_t bytes[WIG_PART_COUNT];
    void *part[WIG_PART_COUNT];
    uint8_t *kind;               // 256 element kinds
    uint32_t *scratch;           // B words
    void *fine;                  // the solve's packed fine level (Dev::coef ... cflag)
    void *vlevels;               // the 16-bit V-cycle's coarse levels
    void *check64;               // wig_capture / wig_cg64's arrays (made on first use)
    uint64_t vlevel_bytes;
    unsigned long long *totals;  // WigRunTotals on the device
    uint32_t *wstate;            // WS_* words on the device (the end of the solve part)
    size_t lds_bytes;            // k_coarseLDS's dynamic LDS
    uint32_t lds_first;          // the first coarse level LDS can hold with all below it
    uint32_t tile[MG_LEVELS + 1];   // tile side asked per level (wig_set_tile)
    uint32_t last_hint;          // iterations its last solve ran, as the host saw (when to check first)
    int device;
    hipStream_t own;
    Batch *solo;                 // the world alone, as wig_fluid_step steps it
};

// Worlds stepped together: the device table of their Dev, the launch
// shapes' workgroup maps, and the stream and timing of their launches.
struct Batch {
    std::vector<WigWorld *> w;
    std::vector<Dev> htab;
    Dev *tab = nullptr;
    uint32_t *maps = nullptr;    // per shape: map (a world index a workgroup), then first (n + 1)
    size_t maps_words = 0;
    uint64_t map_off[SH_COUNT] = {}, first_off[SH_COUNT] = {};
    uint32_t total[SH_COUNT] = {};
    std::vector<uint64_t> sig;   // the worlds and tile sides the maps were built for
    uint32_t *stops = nullptr, *stops_host = nullptr;   // k_stops' words (device, pinned host)
    uint32_t *reports = nullptr;
    // the reports asked for without waiting: two pinned slots of rep_cap
    // worlds' words, the events that mark each landed, the worlds in each,
    // the next ticket and the first still valid (earlier ones belong to the
    // slots' last size or to the worlds before an assign)
    uint32_t *rep_host = nullptr;
    uint32_t rep_cap = 0, rep_n[2] = {};
    hipEvent_t rep_ev[2] = {};
    bool rep_made = false;
    uint64_t rep_next = 0, rep_floor = 0;
    hipStream_t stream = nullptr;
    hipStream_t own = nullptr;   // wig_batch_create's: the stream when none is set (not a world's, which may go first)
    // hipGraphs of the step's phases (run_phase),
    // captured for the launch layout `gsig` and replayed while it holds
    bool graphs = false;
    hipGraphExec_t gx[6] = {};
    Batch *tail = nullptr;       // the worlds still iterating, once few are (batch_step)
    std::vector<uint64_t> gsig;
    int timed = -1;              // kernel id timed, -1 none
    uint32_t ntimed = 0;
    hipEvent_t ev[2 * MAX_TIMED];
    bool ev_made = false;
    int device = 0;              // the device of the worlds it was made for
};

static WorldGrid grid_of(const Batch &b, int sh) {
    WorldGrid g;
    g.tab = b.tab;
    g.map = b.maps + b.map_off[sh];
    g.first = b.maps + b.first_off[sh];
    g.solo = b.w.size() == 1u ? 1u : 0u;
    return g;
}

// A launch over the batch in shape `sh`, bracketed by events when its kernel
// is the one being timed.
template <typename... KArgs, typename... Args>
static void launch_lds(Batch &b, int id, void (*k)(KArgs...), int sh, uint32_t threads, size_t lds, Args... args) {
    if (b.total[sh] == 0u) return;
    bool t = b.timed == id && b.ntimed < MAX_TIMED;
    if (t) (void)hipEventRecord(b.ev[2 * b.ntimed], b.stream);
    hipLaunchKernelGGL(k, dim3(b.total[sh]), dim3(threads), lds, b.stream, grid_of(b, sh), args...);
    if (t) (void)hipEventRecord(b.ev[2 * b.ntimed + 1], b.stream), b.ntimed++;
}
template <typename... KArgs, typename... Args>
static void launch(Batch &b, int id, void (*k)(KArgs...), int sh, uint32_t threads, Args... args) {
    launch_lds(b, id, k, sh, threads, 0, args...);
}

// The world's sizes and parts (js/sim.js:741-786; fluid.wgsl:2228-2248).
static void size_world(WigWorld *w, uint32_t width, uint32_t height, uint32_t coarse) {
    Dev &d = w->d;
    d.W = (int32_t)width;
    d.H = (int32_t)height;
    d.C = width * height;
    d.F = (width + 1u) * height + width * (height + 1u);
    d.bw = ceil_div(width, WIG_BLOCK);
    d.bh = ceil_div(height, WIG_BLOCK);
    d.B = d.bw * d.bh;
    d.cap = (uint32_t)std::ceil((double)d.C * LIQUID_PER_CELL * 1.25);
    uint32_t L = 0;
    d.lw[0] = width;
    d.lh[0] = height;
    for (uint32_t l = 1; l <= MG_LEVELS; l++) {
        d.lw[l] = (width + (1u << l) - 1u) >> l;
        d.lh[l] = (height + (1u << l) - 1u) >> l;
        L = l;
        if (coarse ? l >= coarse : d.lw[l] * d.lh[l] <= (uint32_t)MG_COARSEST) break;
    }
    d.L = L;
    uint64_t at = 2ull * d.C + d.F + 2ull * ((d.F + 3u) / 4u);
    for (uint32_t l = 1; l <= L + 1; l++) {
        d.lstart[l] = at;
        if (l <= L) at += 2ull * d.lw[l] * d.lh[l];
    }
    w->bytes[WIG_PART_CELLS] = 16ull * d.C + 16ull * ((d.C + 1u) / 2u);
    w->bytes[WIG_PART_PARTICLES0] = 16ull + 48ull * d.cap;
    w->bytes[WIG_PART_PARTICLES1] = w->bytes[WIG_PART_PARTICLES0];
    w->bytes[WIG_PART_COUNTERS] = 4ull * (LIQUID_CELL_WORDS * (uint64_t)d.C + 2ull * d.F);
    w->bytes[WIG_PART_FIELD] = 16ull * at;
    w->bytes[WIG_PART_BLOCKS] = 4ull * (5ull * d.B + FLUID_STATS);
    w->bytes[WIG_PART_FLUIDCG] = 16ull * (d.C + d.B + 1u);
    // (then the world's WS_* words, so a snapshot of the part carries them)
    w->bytes[WIG_PART_SOLVE] = sizeof(SolveHead) + 32ull * d.B + 4ull * WS_WORDS;
    w->bytes[WIG_PART_AIR] = 16ull * d.C;
    w->bytes[WIG_PART_AIR_SOURCE] = 4ull * SOURCE_WORDS * d.C;
    w->bytes[WIG_PART_OCCUPANTS] = 4ull * d.C;
    w->bytes[WIG_PART_GRAVITY] = 16ull * ((d.W + 1u) / 2u) * ((d.H + 1u) / 2u);
    w->bytes[WIG_PART_MOTION] = 4ull * WIG_MOTION_WORDS;
    w->bytes[WIG_PART_SWEPT] = 8ull * d.C;
    w->bytes[WIG_PART_PUSHES] = 4ull * WIG_MAX_BODIES * WIG_BODY_PUSH_WORDS;
}


extern "C" {

const char *wig_last_error(void) { return g_err.c_str(); }

int wig_device_ok(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return 0;
    return 1;
}

static int zero(hipStream_t s, void *p, uint64_t bytes) {
    uint64_t words = bytes / 4u;
    if (!words) return WIG_OK;
    uint32_t grid = (uint32_t)std::min<uint64_t>((words + 255u) / 256u, 65535u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, s, (uint32_t *)p, words, 0u);
    HIPCHECK(hipGetLastError());
    return WIG_OK;
}
// The parts of the step's optional inputs (the air's, the people's, the
// gravity pass's, the bodies') are allocated when a caller first uploads one
// or steps with its switch on: most worlds use none, and airSource alone is
// 64 B a cell, which every world a server admits would otherwise allocate
// and clear.  Until then a download reads zeros (what a fresh part holds).
static bool lazyPart(int p) {
    return p == WIG_PART_AIR || p == WIG_PART_AIR_SOURCE || p == WIG_PART_OCCUPANTS || p == WIG_PART_GRAVITY
           || p == WIG_PART_MOTION || p == WIG_PART_SWEPT || p == WIG_PART_PUSHES;
}

static void batch_free(Batch *b) {
    if (!b) return;
    if (b->tail) batch_free(b->tail);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->tab) (void)hipFree(b->tab);
    if (b->maps) (void)hipFree(b->maps);
    if (b->stops) (void)hipFree(b->stops);
    if (b->stops_host) (void)hipHostFree(b->stops_host);
    if (b->reports) (void)hipFree(b->reports);
    if (b->rep_host) (void)hipHostFree(b->rep_host);
    if (b->rep_made)
        for (hipEvent_t &e : b->rep_ev) (void)hipEventDestroy(e);
    for (hipGraphExec_t &x : b->gx)
        if (x) (void)hipGraphExecDestroy(x);
    if (b->ev_made)
        for (int i = 0; i < 2 * MAX_TIMED; i++) (void)hipEventDestroy(b->ev[i]);
    if (b->own) (void)hipStreamDestroy(b->own);
    delete b;
}

static Batch *batch_new(WigWorld *const *worlds, uint32_t n, hipStream_t stream) {
    Batch *b = new Batch();
    b->w.assign(worlds, worlds + n);
    b->device = worlds[0]->device;
    b->htab.resize(n);
    b->stream = stream;
    if (hipMalloc((void **)&b->tab, sizeof(Dev) * n) != hipSuccess || hipMalloc((void **)&b->stops, 8ull * n) != hipSuccess
        || hipHostMalloc((void **)&b->stops_host, 8ull * n, hipHostMallocDefault) != hipSuccess
        || hipMalloc((void **)&b->reports, 4ull * FLUID_STATS * n) != hipSuccess) {
        fail(WIG_E_NOMEM, "batch allocation failed");
        batch_free(b);
        return nullptr;
    }
    return b;
}

// The shapes' workgroup maps for the worlds' sizes and tile sides.
static int batch_maps(Batch &b) {
    std::vector<uint64_t> sig;
    for (WigWorld *w : b.w) {
        sig.push_back((uint64_t)(uintptr_t)w);
        sig.push_back(((uint64_t)(uint32_t)w->d.W << 32) | (uint32_t)w->d.H);   // (a new world may take a freed one's address)
        sig.push_back(w->d.lds_level);
        for (uint32_t l = 0; l <= MG_LEVELS; l++) sig.push_back(w->d.tile[l]);
    }
    if (b.maps && sig == b.sig) return WIG_OK;
    uint32_t n = (uint32_t)b.w.size();
    std::vector<uint32_t> words;
    for (int sh = 0; sh < SH_COUNT; sh++) {
        std::vector<uint32_t> count(n);
        uint64_t natural = 0;
        for (uint32_t i = 0; i < n; i++) natural += count[i] = shape_count(b.w[i]->d, sh);
        // The particle passes' and the CG's four-block passes' grids over a
        // batch: each world's own grid, scaled down together past a few
        // rounds of what the chip holds at once (workgroups that only find
        // their world's range ended still cost a dispatch each).  The
        // 64-thread passes keep theirs: more of them in flight hides more
        // of their latency.
        const uint64_t budget = 8192u;
        if ((sh == SH_BG16 || sh == SH_PG) && natural > budget)
            for (uint32_t i = 0; i < n; i++) count[i] = std::max<uint32_t>(1u, (uint32_t)(count[i] * budget / natural));
        std::vector<uint32_t> first(n + 1u, 0u);
        for (uint32_t i = 0; i < n; i++) first[i + 1u] = first[i] + count[i];
        b.total[sh] = first[n];
        b.map_off[sh] = words.size();
        for (uint32_t i = 0; i < n; i++) words.insert(words.end(), first[i + 1u] - first[i], i);
        b.first_off[sh] = words.size();
        words.insert(words.end(), first.begin(), first.end());
    }
    if (words.size() > b.maps_words) {
        if (b.maps) HIPCHECK(hipFree(b.maps));
        b.maps = nullptr;
        if (hipMalloc((void **)&b.maps, 4ull * words.size()) != hipSuccess) return fail(WIG_E_NOMEM, "batch maps: %zu words", words.size());
        b.maps_words = words.size();
    }
    HIPCHECK(hipMemcpyAsync(b.maps, words.data(), 4ull * words.size(), hipMemcpyHostToDevice, b.stream));
    HIPCHECK(hipStreamSynchronize(b.stream));
    b.sig = sig;
    return WIG_OK;
}

WigWorld *wig_create(uint32_t width, uint32_t height, uint32_t coarse_levels, uint32_t flags) {
    (void)flags;
    if (width < 2u || height < 2u || width > 16384u || height > 16384u || coarse_levels > MG_LEVELS) {
        fail(WIG_E_INVALID, "invalid world size %ux%u (coarse levels %u)", width, height, coarse_levels);
        return nullptr;
    }
    if ((uint64_t)width * height * 5ull >= (1ull << 32) / 12ull) {
        fail(WIG_E_INVALID, "world %ux%u too large for 32-bit particle slots", width, height);
        return nullptr;
    }
    WigWorld *w = new WigWorld();
    size_world(w, width, height, coarse_levels);
    w->vlevel_bytes = 64u;
    for (uint32_t l = 1; l <= w->d.L; l++)   // (and each level's tile flags, for tiles of 8 or more)
        w->vlevel_bytes += 12ull * w->d.lw[l] * w->d.lh[l] + 64u + ceil_div(w->d.lw[l], 8u) * ceil_div(w->d.lh[l], 8u) + 64u
                           + 8ull * w->d.lw[l] * w->d.lh[l] + 64u + 8ull * w->d.lw[l] * w->d.lh[l] + 64u;
    if (hipGetDevice(&w->device) != hipSuccess || hipStreamCreateWithFlags(&w->own, hipStreamNonBlocking) != hipSuccess) {
        fail(WIG_E_DEVICE, "no HIP device or stream");
        delete w;
        return nullptr;
    }
    for (int p = 0; p < WIG_PART_COUNT; p++) {
        if (lazyPart(p)) continue;   // (allocated when first uploaded or stepped with: ensure_part)
        if (hipMalloc(&w->part[p], w->bytes[p]) != hipSuccess) {
            fail(WIG_E_NOMEM, "hipMalloc of part %d (%llu bytes) failed", p, (unsigned long long)w->bytes[p]);
            wig_destroy(w);
            return nullptr;
        }
    }
    if (hipMalloc((void **)&w->kind, 256 + 256 * sizeof(WigElementProps)) != hipSuccess
        || hipMalloc((void **)&w->scratch, 4ull * w->d.B) != hipSuccess
        || hipMalloc(&w->fine, 77ull * 64ull * w->d.B + 256u) != hipSuccess
        || hipMalloc(&w->vlevels, w->vlevel_bytes) != hipSuccess
        || hipMalloc((void **)&w->totals, sizeof(WigRunTotals)) != hipSuccess) {
        fail(WIG_E_NOMEM, "hipMalloc failed");
        wig_destroy(w);
        return nullptr;
    }
    for (int p = 0; p < WIG_PART_COUNT; p++)
        if (w->part[p] && zero(w->own, w->part[p], w->bytes[p]) != WIG_OK) { wig_destroy(w); return nullptr; }
    zero(w->own, w->kind, 256 + 256 * sizeof(WigElementProps));
    zero(w->own, w->fine, 77ull * 64ull * w->d.B + 256u);
    zero(w->own, w->vlevels, w->vlevel_bytes);
    zero(w->own, w->totals, sizeof(WigRunTotals));
    Dev &d = w->d;
    d.totals = w->totals;
    w->wstate = (uint32_t *)((uint8_t *)w->part[WIG_PART_SOLVE] + sizeof(SolveHead) + 32ull * d.B);
    d.wstate = w->wstate;
    // the coarse levels one workgroup's LDS holds (k_coarseLDS): from the
    // first that fits with all below it
    int lds_max = 65536;
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, w->device);
    d.lds_level = d.L;
    for (uint32_t l = d.L; l >= 1; l--) {
        size_t bytes = 0;
        for (uint32_t k = l; k <= d.L; k++) bytes += 4ull * LDS_LEVEL_FLOATS * d.lw[k] * d.lh[k];
        if (bytes > (size_t)lds_max) break;
        d.lds_level = l;
        w->lds_bytes = bytes;
    }
    if (w->lds_bytes == 0) w->lds_bytes = 4ull * LDS_LEVEL_FLOATS * d.lw[d.L] * d.lh[d.L];
    w->lds_first = d.lds_level;
    if (hipFuncSetAttribute((const void *)k_coarseLDS<STORE_F32>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess
        || hipFuncSetAttribute((const void *)k_coarseLDS<STORE_F16>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess
        || hipFuncSetAttribute((const void *)k_coarseLDS<STORE_BF16>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess
        || w->lds_bytes > (size_t)lds_max) {
        fail(WIG_E_DEVICE, "k_coarseLDS cannot take %zu bytes of LDS", w->lds_bytes);
        wig_destroy(w);
        return nullptr;
    }
    for (uint32_t l = 0; l <= MG_LEVELS; l++) w->tile[l] = d.tile[l] = 16u;
    d.wave_cells = 256u;
    d.level_pairs = 1u;
    d.cells = (uint32_t *)w->part[WIG_PART_CELLS];
    d.sorted = (uint32_t *)w->part[WIG_PART_PARTICLES0];
    d.parts = (uint32_t *)w->part[WIG_PART_PARTICLES1];
    d.counters = (uint32_t *)w->part[WIG_PART_COUNTERS];
    d.field = (float *)w->part[WIG_PART_FIELD];
    d.blocks = (uint32_t *)w->part[WIG_PART_BLOCKS];
    d.cg = (float *)w->part[WIG_PART_FLUIDCG];
    d.head = (SolveHead *)w->part[WIG_PART_SOLVE];
    d.bpart = (double *)((uint8_t *)w->part[WIG_PART_SOLVE] + sizeof(SolveHead));
    d.air = (const float4 *)w->part[WIG_PART_AIR];
    d.source = (uint32_t *)w->part[WIG_PART_AIR_SOURCE];
    d.occ = (const uint32_t *)w->part[WIG_PART_OCCUPANTS];
    d.grav = (const float4 *)w->part[WIG_PART_GRAVITY];
    d.motion = (uint32_t *)w->part[WIG_PART_MOTION];
    d.swept = (const uint32_t *)w->part[WIG_PART_SWEPT];
    d.pushes = (uint32_t *)w->part[WIG_PART_PUSHES];
    d.kind = w->kind;
    d.eprops = (const WigElementProps *)(w->kind + 256);   // (the kinds' 256 bytes, then the columns)
    d.scratch = w->scratch;
    {
        // (block-major, 64 cells a block: packIndex)
        uint8_t *f = (uint8_t *)w->fine;
        const uint64_t P = 64ull * d.B;
        d.coef = (float4 *)f;
        d.pa = (float2 *)(f + 16ull * P);
        d.pb = (float2 *)(f + 24ull * P);
        d.r2 = (float2 *)(f + 32ull * P);
        d.cx = (float *)(f + 40ull * P);
        d.cr = (float *)(f + 44ull * P);
        d.cp0 = (float *)(f + 48ull * P);
        d.cp1 = (float *)(f + 52ull * P);
        d.cq = (float *)(f + 56ull * P);
        d.ha = (uint32_t *)(f + 60ull * P);
        d.hb = (uint32_t *)(f + 64ull * P);
        d.hr2 = (uint32_t *)(f + 68ull * P);
        d.cflag = f + 72ull * P;
        d.cb = (float *)(f + 73ull * P);
        // the 16-bit V-cycle's coarse levels: four halves and a float a cell
        uint8_t *v = (uint8_t *)w->vlevels;
        for (uint32_t l = 1; l <= d.L; l++) {
            uint64_t n = (uint64_t)d.lw[l] * d.lh[l];
            d.vlev[l] = (uint16_t *)v;
            d.vrho[l] = (float *)(v + 8ull * n);
            v += 12ull * n + 64u;
        }
        // which tiles of each tiled level hold fluid (k_levelDown, each solve)
        for (uint32_t l = 1; l <= d.L; l++) {
            d.tflag[l] = v;
            v += ceil_div(d.lw[l], 8u) * ceil_div(d.lh[l], 8u) + 64u;
        }
        // and each level's correction after the down pass (k_levelDown to k_levelUp)
        v = (uint8_t *)(((uintptr_t)v + 63u) & ~(uintptr_t)63u);
        for (uint32_t l = 1; l <= d.L; l++) {
            d.pdown[l] = (uint2 *)v;
            v += 8ull * d.lw[l] * d.lh[l] + 64u;
        }
        // and each level's face betas (k_levelDown's first iteration, each solve)
        for (uint32_t l = 1; l <= d.L; l++) {
            d.vbeta[l] = (float2 *)v;
            v += 8ull * d.lw[l] * d.lh[l] + 64u;
        }
    }
    WigWorld *self = w;
    w->solo = batch_new(&self, 1u, w->own);
    if (!w->solo || hipStreamSynchronize(w->own) != hipSuccess) {
        if (w->solo) fail(WIG_E_DEVICE, "stream failed at creation");
        wig_destroy(w);
        return nullptr;
    }
    return w;
}

int wig_destroy(WigWorld *w) {
    if (!w) return WIG_OK;
    if (w->solo) batch_free(w->solo);
    if (w->own) (void)hipStreamSynchronize(w->own);
    for (int p = 0; p < WIG_PART_COUNT; p++)
        if (w->part[p]) (void)hipFree(w->part[p]);
    if (w->kind) (void)hipFree(w->kind);
    if (w->scratch) (void)hipFree(w->scratch);
    if (w->fine) (void)hipFree(w->fine);
    if (w->vlevels) (void)hipFree(w->vlevels);
    if (w->check64) (void)hipFree(w->check64);
    if (w->totals) (void)hipFree(w->totals);
    if (w->own) (void)hipStreamDestroy(w->own);
    delete w;
    return WIG_OK;
}

static hipStream_t stream_of(const WigWorld *w) { return w->solo->stream; }

int wig_set_kinds(WigWorld *w, const uint8_t kinds[256]) {
    if (!w || !kinds) return fail(WIG_E_INVALID, "null argument");
    HIPCHECK(hipMemcpyAsync(w->kind, kinds, 256, hipMemcpyHostToDevice, stream_of(w)));
    HIPCHECK(hipStreamSynchronize(stream_of(w)));
    return WIG_OK;
}

int wig_set_element_props(WigWorld *w, const WigElementProps props[256]) {
    if (!w || !props) return fail(WIG_E_INVALID, "null argument");
    HIPCHECK(hipMemcpyAsync(w->kind + 256, props, 256 * sizeof(WigElementProps), hipMemcpyHostToDevice, stream_of(w)));
    HIPCHECK(hipStreamSynchronize(stream_of(w)));
    return WIG_OK;
}

int wig_layout(const WigWorld *w, WigLayoutInfo *out) {
    if (!w || !out) return fail(WIG_E_INVALID, "null argument");
    memset(out, 0, sizeof *out);
    const Dev &d = w->d;
    out->width = (uint32_t)d.W;
    out->height = (uint32_t)d.H;
    out->cells = d.C;
    out->faces = d.F;
    out->blocks = d.B;
    out->blocks_wide = d.bw;
    out->capacity = d.cap;
    out->levels = d.L;
    for (uint32_t l = 0; l <= d.L; l++) {
        out->level_width[l] = d.lw[l];
        out->level_height[l] = d.lh[l];
    }
    for (int p = 0; p < WIG_PART_COUNT; p++) out->part_bytes[p] = w->bytes[p];
    return WIG_OK;
}

// (a lazy part's memory, zeroed, made now if it is not yet)
static int ensure_part(WigWorld *w, int p) {
    if (w->part[p]) return WIG_OK;
    void *mem = nullptr;
    if (hipMalloc(&mem, w->bytes[p]) != hipSuccess) return fail(WIG_E_NOMEM, "hipMalloc of part %d failed", p);
    w->part[p] = mem;
    int st = zero(stream_of(w), mem, w->bytes[p]);
    if (st < 0) return st;
    HIPCHECK(hipStreamSynchronize(stream_of(w)));   // (cleared before any stream, a batch's too, uses it)
    if (p == WIG_PART_AIR) w->d.air = (const float4 *)mem;
    if (p == WIG_PART_AIR_SOURCE) w->d.source = (uint32_t *)mem;
    if (p == WIG_PART_OCCUPANTS) w->d.occ = (const uint32_t *)mem;
    if (p == WIG_PART_GRAVITY) w->d.grav = (const float4 *)mem;
    if (p == WIG_PART_MOTION) w->d.motion = (uint32_t *)mem;
    if (p == WIG_PART_SWEPT) w->d.swept = (const uint32_t *)mem;
    if (p == WIG_PART_PUSHES) w->d.pushes = (uint32_t *)mem;
    return WIG_OK;
}

int wig_upload(WigWorld *w, int part, const void *src, size_t bytes) {
    if (!w || !src || part < 0 || part >= WIG_PART_COUNT) return fail(WIG_E_INVALID, "bad part %d", part);
    if (bytes != w->bytes[part]) return fail(WIG_E_INVALID, "part %d is %llu bytes, not %zu", part, (unsigned long long)w->bytes[part], bytes);
    int st = ensure_part(w, part);
    if (st < 0) return st;
    HIPCHECK(hipMemcpyAsync(w->part[part], src, bytes, hipMemcpyHostToDevice, stream_of(w)));
    HIPCHECK(hipStreamSynchronize(stream_of(w)));
    return WIG_OK;
}

int wig_download(WigWorld *w, int part, void *dst, size_t bytes) {
    if (!w || !dst || part < 0 || part >= WIG_PART_COUNT) return fail(WIG_E_INVALID, "bad part %d", part);
    if (bytes != w->bytes[part]) return fail(WIG_E_INVALID, "part %d is %llu bytes, not %zu", part, (unsigned long long)w->bytes[part], bytes);
    if (!w->part[part]) {            // (a lazy part not made yet: zeros)
        memset(dst, 0, bytes);
        return WIG_OK;
    }
    HIPCHECK(hipMemcpyAsync(dst, w->part[part], bytes, hipMemcpyDeviceToHost, stream_of(w)));
    HIPCHECK(hipStreamSynchronize(stream_of(w)));
    return WIG_OK;
}

int wig_device_ptr(WigWorld *w, int part, void **ptr) {
    if (!w || !ptr || part < 0 || part >= WIG_PART_COUNT) return fail(WIG_E_INVALID, "bad part %d", part);
    int st = ensure_part(w, part);
    if (st < 0) return st;
    *ptr = w->part[part];
    return WIG_OK;
}

// Work enqueued on `to` from now on waits for what `from` holds so far (an
// event recorded on `from`); nothing when they are the same stream or
// `from` has finished all its work (hipStreamQuery: the common case, a
// serving loop's streams drained by its reports, where an event's record
// and wait had cost its rounds up to 15 ms at p99).  The event is the
// calling thread's own, made once (per device) and recorded again each
// time: a wait takes the record made before it, so a later record does not
// move it.
int wig_order_streams(hipStream_t from, hipStream_t to) {
    if (from == to) return WIG_OK;
    const hipError_t idle = hipStreamQuery(from);
    if (idle == hipSuccess) return WIG_OK;
    if (idle == hipErrorNotReady) (void)hipGetLastError();   // (not an error: kept from the next call's check)
    static thread_local hipEvent_t ev = nullptr;
    static thread_local int ev_device = -1;
    int device = 0;
    HIPCHECK(hipGetDevice(&device));
    if (!ev || ev_device != device) {
        HIPCHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));   // (one a thread and device, kept)
        ev_device = device;
    }
    hipError_t e = hipEventRecord(ev, from);
    if (e == hipSuccess) e = hipStreamWaitEvent(to, ev, 0);
    if (e != hipSuccess) return fail(WIG_E_DEVICE, "ordering the new stream after the old: %s", hipGetErrorString(e));
    return WIG_OK;
}

int wig_set_stream(WigWorld *w, void *stream) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    hipStream_t next = stream ? (hipStream_t)stream : w->own;
    int st = wig_order_streams(w->solo->stream, next);
    if (st < 0) return st;
    w->solo->stream = next;
    return WIG_OK;
}

int wig_sync(WigWorld *w) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    HIPCHECK(hipStreamSynchronize(stream_of(w)));
    return WIG_OK;
}

// the worlds' reports gathered into b.reports, behind the enqueued work
static int enqueue_reports(Batch &b) {
    uint32_t n = (uint32_t)b.w.size();
    int st = batch_maps(b);
    if (st < 0) return st;
    for (uint32_t i = 0; i < n; i++) b.htab[i] = b.w[i]->d;
    HIPCHECK(hipMemcpyAsync(b.tab, b.htab.data(), sizeof(Dev) * n, hipMemcpyHostToDevice, b.stream));
    hipLaunchKernelGGL(k_reports, dim3(n), dim3(64), 0, b.stream, grid_of(b, SH_ONE), b.reports);
    HIPCHECK(hipGetLastError());
    return WIG_OK;
}

static int batch_report(Batch &b, uint32_t *stats) {
    int st = enqueue_reports(b);
    if (st < 0) return st;
    HIPCHECK(hipMemcpyAsync(stats, b.reports, 4ull * FLUID_STATS * b.w.size(), hipMemcpyDeviceToHost, b.stream));
    HIPCHECK(hipStreamSynchronize(b.stream));
    return WIG_OK;
}

static int batch_report_async(Batch &b, uint64_t *ticket) {
    if (!ticket) return fail(WIG_E_INVALID, "null ticket");
    const uint32_t n = (uint32_t)b.w.size();
    if (!b.rep_made) {
        for (hipEvent_t &e : b.rep_ev) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b.rep_made = true;
    }
    if (n > b.rep_cap) {
        // (the slots may still be filling: both landed before they go; an
        // event never recorded has nothing to wait for)
        for (hipEvent_t &e : b.rep_ev) HIPCHECK(hipEventSynchronize(e));
        uint32_t *h = nullptr;
        if (hipHostMalloc((void **)&h, 2ull * 4ull * FLUID_STATS * n, hipHostMallocDefault) != hipSuccess)
            return fail(WIG_E_NOMEM, "report slots: pinned allocation failed");
        if (b.rep_host) HIPCHECK(hipHostFree(b.rep_host));
        b.rep_host = h;
        b.rep_cap = n;
        b.rep_floor = b.rep_next;
    }
    int st = enqueue_reports(b);
    if (st < 0) return st;
    const uint32_t slot = (uint32_t)(b.rep_next & 1u);
    HIPCHECK(hipMemcpyAsync(b.rep_host + (uint64_t)slot * FLUID_STATS * b.rep_cap, b.reports, 4ull * FLUID_STATS * n,
                            hipMemcpyDeviceToHost, b.stream));
    HIPCHECK(hipEventRecord(b.rep_ev[slot], b.stream));
    b.rep_n[slot] = n;
    *ticket = b.rep_next++;
    return WIG_OK;
}

static int batch_report_take(Batch &b, uint64_t ticket, uint32_t *stats, int wait) {
    if (!stats) return fail(WIG_E_INVALID, "null stats");
    if (!b.rep_made || ticket < b.rep_floor || ticket >= b.rep_next || ticket + 2u < b.rep_next)
        return fail(WIG_E_INVALID, "report %llu is not one of the last two asked for", (unsigned long long)ticket);
    const uint32_t slot = (uint32_t)(ticket & 1u);
    if (wait) {
        HIPCHECK(hipEventSynchronize(b.rep_ev[slot]));
    } else {
        const hipError_t e = hipEventQuery(b.rep_ev[slot]);
        if (e == hipErrorNotReady) {
            (void)hipGetLastError();   // (not an error: kept from the next launch's check)
            return WIG_NOT_READY;
        }
        HIPCHECK(e);
    }
    memcpy(stats, b.rep_host + (uint64_t)slot * FLUID_STATS * b.rep_cap, 4ull * FLUID_STATS * b.rep_n[slot]);
    return WIG_OK;
}

int wig_report_async(WigWorld *w, uint64_t *ticket) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    return batch_report_async(*w->solo, ticket);
}

int wig_report_take(WigWorld *w, uint64_t ticket, uint32_t stats[WIG_STATS], int wait) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    return batch_report_take(*w->solo, ticket, stats, wait);
}

int wig_report(WigWorld *w, uint32_t stats[WIG_STATS]) {
    if (!w || !stats) return fail(WIG_E_INVALID, "null argument");
    return batch_report(*w->solo, stats);
}

int wig_run_totals(WigWorld *w, WigRunTotals *out) {
    if (!w || !out) return fail(WIG_E_INVALID, "null argument");
    HIPCHECK(hipMemcpyAsync(out, w->totals, sizeof *out, hipMemcpyDeviceToHost, stream_of(w)));
    HIPCHECK(hipStreamSynchronize(stream_of(w)));
    return WIG_OK;
}

static int batch_time_kernel(Batch &b, int id) {
    if (id > KID_COUNT) return fail(WIG_E_INVALID, "bad kernel id %d", id);
    if (!b.ev_made) {
        for (int i = 0; i < 2 * MAX_TIMED; i++) HIPCHECK(hipEventCreate(&b.ev[i]));
        b.ev_made = true;
    }
    b.timed = id < 0 ? -1 : id;
    b.ntimed = 0;
    return WIG_OK;
}
static int batch_kernel_time(Batch &b, double *ms, uint32_t *launches) {
    if (!ms || !launches) return fail(WIG_E_INVALID, "null argument");
    HIPCHECK(hipStreamSynchronize(b.stream));
    double total = 0.0;
    for (uint32_t i = 0; i < b.ntimed; i++) {
        float t = 0.0f;
        HIPCHECK(hipEventElapsedTime(&t, b.ev[2 * i], b.ev[2 * i + 1]));
        total += t;
    }
    *ms = total;
    *launches = b.ntimed;
    b.ntimed = 0;
    return WIG_OK;
}

int wig_time_kernel(WigWorld *w, int id) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    return batch_time_kernel(*w->solo, id);
}

int wig_kernel_time(WigWorld *w, double *ms, uint32_t *launches) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    return batch_kernel_time(*w->solo, ms, launches);
}

int wig_set_tile(WigWorld *w, uint32_t level, uint32_t side) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    if (level < 1u || level >= w->d.lds_level) return fail(WIG_E_INVALID, "level %u is not swept by tiles (1 .. %u)", level, w->d.lds_level - 1u);
    if (side != 8u && side != 16u && side != 32u) return fail(WIG_E_INVALID, "tile side %u: 8, 16 or 32", side);
    w->tile[level] = side;
    return WIG_OK;
}

int wig_set_lds_level(WigWorld *w, uint32_t level) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    Dev &d = w->d;
    if (level < w->lds_first || level > d.L)
        return fail(WIG_E_INVALID, "level %u: the one-workgroup kernel can start at levels %u .. %u", level, w->lds_first, d.L);
    d.lds_level = level;
    w->lds_bytes = 0;
    for (uint32_t k = level; k <= d.L; k++) w->lds_bytes += 4ull * LDS_LEVEL_FLOATS * d.lw[k] * d.lh[k];
    return WIG_OK;
}

int wig_set_coarse_wave(WigWorld *w, uint32_t cells) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    w->d.wave_cells = cells;
    return WIG_OK;
}

const char *wig_kernel_name(int id) { return id >= 0 && id < KID_COUNT ? g_kernel_names[id] : (id == KID_COUNT ? "solve" : ""); }
int wig_kernel_count(void) { return KID_COUNT + 1; }

int wig_solve_stats(WigWorld *w, WigSolveStats *out) {
    if (!w || !out) return fail(WIG_E_INVALID, "null argument");
    HIPCHECK(hipMemcpyAsync(out, w->part[WIG_PART_SOLVE], sizeof *out, hipMemcpyDeviceToHost, stream_of(w)));
    HIPCHECK(hipStreamSynchronize(stream_of(w)));
    return WIG_OK;
}

// The step's words of world w's Dev, from its Params and controls.
static int step_words(WigWorld *w, const WigParams *P, const WigStepCtl *ctl, Dev &d) {
    const bool need[WIG_PART_COUNT] = {false, false, false, false, false, false, false, false,
                                        ctl->air != 0u, ctl->air != 0u, ctl->occupants != 0u, ctl->gravity != 0u,
                                        ctl->bodies != 0u, ctl->bodies != 0u, ctl->bodies != 0u};
    static_assert(WIG_PART_COUNT == 15, "the parts a step's switches need");
    for (int p = 0; p < WIG_PART_COUNT; p++)
        if (need[p]) {
            int st = ensure_part(w, p);
            if (st < 0) return st;
        }
    d = w->d;
    if (P->size[0] != (uint32_t)d.W || P->size[1] != (uint32_t)d.H)
        return fail(WIG_E_INVALID, "params size %ux%u is not the world's %dx%d", P->size[0], P->size[1], d.W, d.H);
    if (ctl->surface && ctl->residual_every)
        return fail(WIG_E_INVALID, "residual replacement takes the open cells' pressure as 0: not with surface tension");
    if (ctl->bodies && ctl->residual_every)
        return fail(WIG_E_INVALID, "residual replacement leaves out the bodies' coupling: not with bodies");
    if (ctl->bodies_run && !ctl->bodies) return fail(WIG_E_INVALID, "bodies_run needs the bodies' parts (ctl.bodies)");
    if (ctl->fluid_cycles > WIG_MAX_CYCLES || ctl->sweep_pairs > 64u || ctl->coarse_sweeps > 4096u || ctl->tol_mode > 1u
        || ctl->level_pairs > MAX_LEVEL_PAIRS || ctl->vcycle_storage > WIG_STORE_BF16 || ctl->residual_every > WIG_MAX_CYCLES)
        return fail(WIG_E_INVALID, "solver control out of range");
    if (ctl->pre_relax != 1u || ctl->post_relax != 1u)
        return fail(WIG_E_INVALID, "one relaxation before and one after the coarse levels (PRE_RELAX = POST_RELAX = 1) is what is built");
    d.frame = P->frame;
    d.seed = P->seed;
    d.earthG = P->earthG;
    d.full = ctl->full ? 1u : 0u;
    d.fluid_cycles = ctl->fluid_cycles;
    d.shifted_first = ctl->shifted_first ? 1u : 0u;
    d.first_chunk = ctl->first_chunk;
    d.chunk = ctl->chunk;
    d.sweep_pairs = ctl->sweep_pairs;
    d.coarse_sweeps = ctl->coarse_sweeps;
    d.tol_mode = ctl->tol_mode;
    d.tol = ctl->tol;
    d.level_pairs = ctl->level_pairs ? ctl->level_pairs : 1u;
    d.store = ctl->vcycle_storage;
    d.replace_every = ctl->residual_every;
    d.air_on = ctl->air ? 1u : 0u;
    d.occ_on = ctl->occupants ? 1u : 0u;
    d.grav_on = ctl->gravity ? 1u : 0u;
    d.visc_on = ctl->viscous ? 1u : 0u;
    d.surf_on = ctl->surface ? 1u : 0u;
    d.body_on = ctl->bodies ? 1u : 0u;
    d.bodies_run = ctl->bodies_run ? 1u : 0u;
    d.surface_drag = ctl->surface_drag;
    d.dt = P->dt;
    d.dx = P->dx;
    for (uint32_t l = 0; l <= MG_LEVELS; l++) d.tile[l] = fitTile(w->tile[l], d.level_pairs);
    w->d.level_pairs = d.level_pairs;
    for (uint32_t l = 0; l <= MG_LEVELS; l++) w->d.tile[l] = d.tile[l];
    return WIG_OK;
}

// What a step launches, for the batch's worlds and their controls.
struct StepPlan {
    bool any_full = false;
    bool replace = false;    // a full world replaces its residual (k_cgTrue in the iteration)
    bool viscous = false;    // a full world runs viscosity's sweeps
    bool surface = false;    // a full world runs the surface pass
    bool bodies = false;     // a full world has bodies (the bodies' passes)
    bool bodies_run = false; // ... and hands them what the liquid did (rigidify, bodyPush)
    uint32_t stores = 0;     // the V-cycle storages (1 << STORE_*) of the full worlds
    uint32_t max_cycles = 0, first_check = NO_STOP, chunk = NO_STOP, max_level = 0;
    size_t coarse_lds = 0, down_lds[MG_LEVELS + 1] = {}, up_lds[MG_LEVELS + 1] = {};
};

// The step's phases: the head every world runs (the idle path's passes,
// and the full path's first), the full path up to the iterations, one
// iteration, and the rest.
static void enqueue_head(Batch &b) {
    launch(b, KID_sync, k_sync, SH_WORLD, 64);
    launch(b, KID_syncParticles, k_syncParticles, SH_PG, 256);
    launch(b, KID_syncSeed, k_syncSeed, SH_WORLD, 64);
    launch(b, KID_compactMarks, k_compactMarks, SH_B256, 256);
    launch(b, KID_compact, k_compact, SH_ONE, 1024);
    launch(b, KID_retire, k_retire, SH_BG, 64);
}
#define VISCOSITY_ITERATIONS 8   // js/sim.js:71
static void enqueue_pre(Batch &b, const StepPlan &pl) {
    launch(b, KID_clear, k_clear, SH_BG, 64);
    launch(b, KID_gather, k_gather, SH_PG, 256);
    launch(b, KID_localJoin, k_localJoin, SH_BG, 64, 0);
    launch(b, KID_edgeJoin, k_edgeJoin, SH_BG, 64, 0);
    launch(b, KID_sealAnchor, k_sealAnchor, SH_BG, 64);
    launch(b, KID_classify, k_classify, SH_BG, 64);
    if (pl.surface) launch(b, KID_surface, k_surface, SH_BG, 64);
    if (pl.bodies) {
        launch(b, KID_bodyStart, k_bodyStart, SH_BG, 64);
        launch(b, KID_bodyLoads, k_bodyLoads, SH_BG, 64);
    }
    launch(b, KID_setFaces, k_setFaces, SH_BG, 64);
    // viscosity's sweeps before the projection; unpress whether they ran or
    // not (js/sim.js:2505-2514)
    for (int i = 0; pl.viscous && i < VISCOSITY_ITERATIONS; i++) {
        launch(b, KID_viscosity, k_viscosity, SH_BG, 64, 0);
        launch(b, KID_viscosity, k_viscosity, SH_BG, 64, 1);
    }
    launch(b, KID_viscosity, k_viscosity, SH_BG, 64, 2);
    // the liquid's drag on the bodies, and their faces moving with them again
    // (js/sim.js:2515-2520)
    if (pl.bodies && pl.viscous) {
        launch(b, KID_bodyDrag, k_bodyDrag, SH_BG, 64);
        launch(b, KID_bodyFaces, k_bodyFaces, SH_BG, 64);
    }
    launch(b, KID_divergence, k_divergence, SH_BG, 64);
}
extern "C++" {
template <int S> static void enqueue_start_store(Batch &b) {
    launch(b, KID_cgPrepare, k_cgPrepare<S>, SH_BG, 64);
    launch(b, KID_levelRho, k_levelRho<S>, SH_LVL1, 64);
}
template <int S> static void enqueue_iteration_store(Batch &b, const StepPlan &pl, int part) {
    if (part == 0) {
        launch(b, KID_relax, k_relax<S>, SH_RELAX, 256, 0);
        launch(b, KID_restrict, k_restrict<S>, SH_LVL1, 64);
        for (uint32_t l = 1; l <= pl.max_level; l++)   // level l's sweeps, and restrictLevel into l + 1
            launch_lds(b, KID_levelDown, k_levelDown<S>, SH_TILE + (int)l - 1, 256, pl.down_lds[l], l);
        launch_lds(b, KID_coarseLDS, k_coarseLDS<S>, SH_ONE, 1024, pl.coarse_lds);
        for (uint32_t l = pl.max_level; l >= 1; l--)
            launch_lds(b, KID_levelUp, k_levelUp<S>, SH_TILE + (int)l - 1, 256, pl.up_lds[l], l);
        launch(b, KID_relax, k_relax<S>, SH_RELAX, 256, 1);
        launch(b, KID_cgRho, k_cgRho<S>, SH_BG16, 256);
    } else if (part == 1) {
        if (pl.bodies) launch(b, KID_cgBodyP, k_cgBodyP<S>, SH_BG, 64);
        launch(b, KID_cgApply, k_cgApply<S>, SH_BG16, 256);
    } else {
        launch(b, KID_cgUpdate, k_cgUpdate<S>, SH_BG16, 256);
        if (pl.replace) launch(b, KID_cgTrue, k_cgTrue<S>, SH_BG16, 256);
    }
}
}  // extern "C++"
static void enqueue_solve_start(Batch &b, const StepPlan &pl) {
    launch(b, KID_cgStart, k_cgStart, SH_BG, 64);
    launch(b, KID_localJoin, k_localJoin, SH_BG, 64, 1);
    launch(b, KID_edgeJoin, k_edgeJoin, SH_BG, 64, 1);
    launch(b, KID_cgAnchor, k_cgAnchor, SH_BG, 64);
    if (pl.bodies) launch(b, KID_cgBodyX, k_cgBodyX, SH_BG, 64);
    launch(b, KID_cgResidual, k_cgResidual, SH_BG4, 256);
    launch(b, KID_sumNorm, k_sumNorm, SH_ONE, 1024);
    if (pl.stores & 1u) enqueue_start_store<STORE_F32>(b);
    if (pl.stores & 2u) enqueue_start_store<STORE_F16>(b);
    if (pl.stores & 4u) enqueue_start_store<STORE_BF16>(b);
}
static void enqueue_iteration(Batch &b, const StepPlan &pl) {
    auto stores = [&](int part) {
        if (pl.stores & 1u) enqueue_iteration_store<STORE_F32>(b, pl, part);
        if (pl.stores & 2u) enqueue_iteration_store<STORE_F16>(b, pl, part);
        if (pl.stores & 4u) enqueue_iteration_store<STORE_BF16>(b, pl, part);
    };
    stores(0);
    launch(b, KID_cgBeta, k_cgBeta, SH_ONE, 1024);
    stores(1);
    launch(b, KID_cgAlpha, k_cgAlpha, SH_ONE, 1024);
    stores(2);
    launch(b, KID_cgCheck, k_cgCheck, SH_ONE, 1024);
}
static void enqueue_post(Batch &b, const StepPlan &pl) {
    if (pl.bodies) launch(b, KID_cgBodyF, k_cgBodyF, SH_BG, 64);
    launch(b, KID_project, k_project, SH_BG, 64);
    // what the liquid did to the bodies in it (js/sim.js:2594-2599)
    if (pl.bodies_run) {
        launch(b, KID_rigidify, k_rigidify, SH_BG, 64);
        launch(b, KID_bodyPush, k_bodyPush, SH_ONE, 1024);
    }
    launch(b, KID_extrapolate, k_extrapolate, SH_BG, 64, 1);
    launch(b, KID_extrapolate, k_extrapolate, SH_BG, 64, 2);
    launch(b, KID_advect, k_advect, SH_PG, 256);
    launch(b, KID_sortCount, k_sortCount, SH_PG, 256);
    launch(b, KID_scanBlocks, k_scanBlocks, SH_BG, 64);
    launch(b, KID_scanTotals, k_scanTotals, SH_ONE, 256);
    launch(b, KID_scanFinish, k_scanFinish, SH_BG, 64);
    launch(b, KID_scatter, k_scatter, SH_PG, 256);
    launch(b, KID_rasterClaims, k_rasterClaims, SH_BG, 64);
    launch(b, KID_raster, k_raster, SH_BG, 64);
    launch(b, KID_unpaint, k_unpaint, SH_BG, 64);
}

// Phase `ph` (0 head, 1 pre, 2 the solve's start, 3 an iteration, 4
// cgFinish, 5 post) enqueued, or, with graphs on, its graph launched
// (captured first when the launch layout changed).
enum { PH_HEAD, PH_PRE, PH_START, PH_ITER, PH_FINISH, PH_POST };
static int run_phase(Batch &b, int ph, const StepPlan &pl) {
    auto enqueue = [&]() {
        switch (ph) {
        case PH_HEAD: enqueue_head(b); break;
        case PH_PRE: enqueue_pre(b, pl); break;
        case PH_START: enqueue_solve_start(b, pl); break;
        case PH_ITER: enqueue_iteration(b, pl); break;
        case PH_FINISH:
            if (pl.stores & 1u) launch(b, KID_cgFinish, k_cgFinish<STORE_F32>, SH_BG, 64);
            if (pl.stores & 2u) launch(b, KID_cgFinish, k_cgFinish<STORE_F16>, SH_BG, 64);
            if (pl.stores & 4u) launch(b, KID_cgFinish, k_cgFinish<STORE_BF16>, SH_BG, 64);
            break;
        default: enqueue_post(b, pl); break;
        }
    };
    if (!b.graphs || b.timed >= 0) {
        enqueue();
        HIPCHECK(hipGetLastError());
        return WIG_OK;
    }
    if (!b.gx[ph]) {
        hipGraph_t graph = nullptr;
        HIPCHECK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
        enqueue();
        HIPCHECK(hipStreamEndCapture(b.stream, &graph));
        hipError_t e = hipGraphInstantiate(&b.gx[ph], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(WIG_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    HIPCHECK(hipGraphLaunch(b.gx[ph], b.stream));
    return WIG_OK;
}

static int batch_assign(Batch *b, WigWorld *const *worlds, uint32_t n) {
    if (!b || !worlds || n == 0u) return fail(WIG_E_INVALID, "a batch needs worlds");
    for (uint32_t i = 0; i < n; i++)
        if (!worlds[i] || worlds[i]->device != b->device) return fail(WIG_E_INVALID, "world %u is null or on another device", i);   // (b's own worlds may be gone)
    if (n > b->htab.size()) {
        Dev *tab = nullptr;
        uint32_t *stops = nullptr, *stops_host = nullptr, *reports = nullptr;
        if (hipMalloc((void **)&tab, sizeof(Dev) * n) != hipSuccess || hipMalloc((void **)&stops, 8ull * n) != hipSuccess
            || hipHostMalloc((void **)&stops_host, 8ull * n, hipHostMallocDefault) != hipSuccess
            || hipMalloc((void **)&reports, 4ull * FLUID_STATS * n) != hipSuccess) {
            // (what was made before the failure, freed; the batch keeps its own)
            if (tab) (void)hipFree(tab);
            if (stops) (void)hipFree(stops);
            if (stops_host) (void)hipHostFree(stops_host);
            if (reports) (void)hipFree(reports);
            return fail(WIG_E_NOMEM, "batch allocation failed");
        }
        HIPCHECK(hipStreamSynchronize(b->stream));
        HIPCHECK(hipFree(b->tab));
        HIPCHECK(hipFree(b->stops));
        HIPCHECK(hipHostFree(b->stops_host));
        HIPCHECK(hipFree(b->reports));
        b->tab = tab;
        b->stops = stops;
        b->stops_host = stops_host;
        b->reports = reports;
    }
    b->w.assign(worlds, worlds + n);
    b->htab.resize(std::max<size_t>(b->htab.size(), n));
    b->rep_floor = b->rep_next;   // (the last reports were of the worlds before)
    b->sig.clear();          // the maps are built again at the next step
    return WIG_OK;
}

// The plan of a batch's iterations: which storages, levels and LDS sizes.
static void plan_iterations(const Batch &b, StepPlan &pl) {
    pl.stores = 0;
    pl.replace = false;
    pl.max_level = 0;
    pl.coarse_lds = 0;
    for (uint32_t l = 0; l <= MG_LEVELS; l++) pl.down_lds[l] = pl.up_lds[l] = 0;
    for (uint32_t i = 0; i < (uint32_t)b.w.size(); i++) {
        const Dev &d = b.htab[i];
        if (!d.full) continue;
        pl.stores |= 1u << d.store;
        pl.replace = pl.replace || d.replace_every != 0u;
        pl.coarse_lds = std::max(pl.coarse_lds, b.w[i]->lds_bytes);
        for (uint32_t l = 1; l < d.lds_level; l++) {
            pl.max_level = std::max(pl.max_level, l);
            pl.down_lds[l] = std::max(pl.down_lds[l], downLds(d.level_pairs, d.tile[l]));
            pl.up_lds[l] = std::max(pl.up_lds[l], upLds(d.level_pairs, d.tile[l]));
        }
    }
}

// One liquid step of every world of the batch (encodeFluid, js/sim.js:2466-2612).
static int batch_step(Batch &b, const WigParams *P, const WigStepCtl *ctl) {
    uint32_t n = (uint32_t)b.w.size();
    bool timeSolve = b.timed == KID_COUNT && b.ntimed < MAX_TIMED;
    StepPlan pl;
    for (uint32_t i = 0; i < n; i++) {
        int st = step_words(b.w[i], &P[i], &ctl[i], b.htab[i]);
        if (st < 0) return st;
    }
    int st = batch_maps(b);
    if (st < 0) return st;
    for (uint32_t i = 0; i < n; i++) {
        const Dev &d = b.htab[i];
        if (!d.full) continue;
        pl.any_full = true;
        pl.stores |= 1u << d.store;
        pl.replace = pl.replace || d.replace_every != 0u;
        pl.viscous = pl.viscous || d.visc_on != 0u;
        pl.surface = pl.surface || d.surf_on != 0u;
        pl.bodies = pl.bodies || d.body_on != 0u;
        pl.bodies_run = pl.bodies_run || d.bodies_run != 0u;
        pl.max_cycles = std::max(pl.max_cycles, d.fluid_cycles);
        if (d.first_chunk != 0u) {
            uint32_t f = d.first_chunk == WIG_CHUNK_AUTO ? std::max(b.w[i]->last_hint, 1u) : d.first_chunk;
            pl.first_check = std::min(pl.first_check, f);
            pl.chunk = std::min(pl.chunk, d.chunk ? d.chunk : 1u);
        }
        pl.coarse_lds = std::max(pl.coarse_lds, b.w[i]->lds_bytes);
        for (uint32_t l = 1; l < d.lds_level; l++) {
            pl.max_level = std::max(pl.max_level, l);
            pl.down_lds[l] = std::max(pl.down_lds[l], downLds(d.level_pairs, d.tile[l]));
            pl.up_lds[l] = std::max(pl.up_lds[l], upLds(d.level_pairs, d.tile[l]));
        }
    }
    if (b.graphs) {
        // the graphs hold the launch layout: maps, LDS sizes, levels, stream
        std::vector<uint64_t> sig(b.sig.begin(), b.sig.end());
        sig.push_back((uint64_t)(uintptr_t)b.maps);
        sig.push_back((uint64_t)(uintptr_t)b.tab);
        sig.push_back(pl.coarse_lds);
        sig.push_back(pl.stores);
        sig.push_back(pl.replace);
        sig.push_back(pl.viscous);
        sig.push_back(pl.surface);
        sig.push_back(pl.bodies);
        sig.push_back(pl.bodies_run);
        sig.push_back(pl.max_level);
        for (uint32_t l = 0; l <= MG_LEVELS; l++) sig.push_back(pl.down_lds[l]), sig.push_back(pl.up_lds[l]);
        sig.push_back((uint64_t)(uintptr_t)b.stream);
        if (sig != b.gsig) {
            for (hipGraphExec_t &x : b.gx) {
                if (x) HIPCHECK(hipGraphExecDestroy(x));
                x = nullptr;
            }
            b.gsig = sig;
        }
    }
    HIPCHECK(hipMemcpyAsync(b.tab, b.htab.data(), sizeof(Dev) * n, hipMemcpyHostToDevice, b.stream));
    hipStream_t s = b.stream;
    if ((st = run_phase(b, PH_HEAD, pl)) < 0) return st;
    if (!pl.any_full) return WIG_OK;
    if ((st = run_phase(b, PH_PRE, pl)) < 0) return st;
    // (the solve, cgStart .. cgFinish, timed as one: the pseudo-kernel "solve")
    if (timeSolve) HIPCHECK(hipEventRecord(b.ev[2 * b.ntimed], s));
    if ((st = run_phase(b, PH_START, pl)) < 0) return st;
    // iterations: each world runs the ones its own host would have encoded
    // (encoded() in the kernels); the batch checks at its checkpoints
    // whether every world has stopped
    // Once a checkpoint finds a quarter of the worlds or fewer still
    // iterating, the rest of the iterations run on a batch of those alone
    // (`cur`), so the finished ones' workgroups are not launched to exit.
    uint32_t next_check = pl.first_check;
    Batch *cur = &b;
    StepPlan cpl = pl;
    for (uint32_t cycle = 0; cycle < pl.max_cycles; cycle++) {
        if (cycle == next_check) {
            uint32_t m = (uint32_t)cur->w.size();
            hipLaunchKernelGGL(k_stops, dim3(m), dim3(64), 0, s, grid_of(*cur, SH_ONE), cur->stops);
            HIPCHECK(hipMemcpyAsync(cur->stops_host, cur->stops, 8ull * m, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            std::vector<uint32_t> going;
            for (uint32_t i = 0; i < m; i++) {
                const Dev &d = cur->htab[i];
                bool done = !d.full || d.fluid_cycles <= cycle || cur->stops_host[2 * i] <= cycle;
                if (done && d.full && cur->stops_host[2 * i] < d.fluid_cycles) cur->w[i]->last_hint = cur->stops_host[2 * i + 1];
                if (!done) going.push_back(i);
            }
            if (going.empty()) break;
            if (cur == &b && (b.timed < 0 || b.timed == KID_COUNT) && 4u * going.size() <= m && going.size() < m) {   // (the solve timed as a whole keeps the tail)
                std::vector<WigWorld *> ws;
                for (uint32_t i : going) ws.push_back(b.w[i]);
                if (!b.tail) {
                    b.tail = batch_new(ws.data(), (uint32_t)ws.size(), s);
                    if (!b.tail) return WIG_E_NOMEM;
                } else if ((st = batch_assign(b.tail, ws.data(), (uint32_t)ws.size())) < 0) {
                    return st;
                }
                Batch &t = *b.tail;
                t.stream = s;
                t.graphs = false;
                for (uint32_t k = 0; k < (uint32_t)going.size(); k++) t.htab[k] = b.htab[going[k]];
                if ((st = batch_maps(t)) < 0) return st;
                HIPCHECK(hipMemcpyAsync(t.tab, t.htab.data(), sizeof(Dev) * going.size(), hipMemcpyHostToDevice, s));
                plan_iterations(t, cpl);
                cur = &t;
            }
            next_check = cycle + pl.chunk;
        }
        if ((st = run_phase(*cur, PH_ITER, cpl)) < 0) return st;
    }
    if ((st = run_phase(b, PH_FINISH, pl)) < 0) return st;
    if (timeSolve) {
        HIPCHECK(hipEventRecord(b.ev[2 * b.ntimed + 1], s));
        b.ntimed++;
    }
    if ((st = run_phase(b, PH_POST, pl)) < 0) return st;
    return WIG_OK;
}

int wig_fluid_step(WigWorld *w, const WigParams *P, const WigStepCtl *ctl) {
    if (!w || !P || !ctl) return fail(WIG_E_INVALID, "null argument");
    return batch_step(*w->solo, P, ctl);
}

WigBatch *wig_batch_create(WigWorld *const *worlds, uint32_t n) {
    if (!worlds || n == 0u) {
        fail(WIG_E_INVALID, "a batch needs worlds");
        return nullptr;
    }
    for (uint32_t i = 0; i < n; i++)
        if (!worlds[i] || worlds[i]->device != worlds[0]->device) {
            fail(WIG_E_INVALID, "world %u is null or on another device", i);
            return nullptr;
        }
    hipStream_t own = nullptr;
    if (hipStreamCreate(&own) != hipSuccess) {
        fail(WIG_E_DEVICE, "hipStreamCreate failed");
        return nullptr;
    }
    Batch *b = batch_new(worlds, n, own);
    if (!b) {
        (void)hipStreamDestroy(own);
        return nullptr;
    }
    b->own = own;
    return (WigBatch *)b;
}

int wig_batch_assign(WigBatch *bb, WigWorld *const *worlds, uint32_t n) {
    return batch_assign((Batch *)bb, worlds, n);
}

int wig_batch_destroy(WigBatch *b) {
    batch_free((Batch *)b);
    return WIG_OK;
}

int wig_batch_set_stream(WigBatch *b, void *stream) {
    if (!b) return fail(WIG_E_INVALID, "null batch");
    Batch *B = (Batch *)b;
    hipStream_t next = stream ? (hipStream_t)stream : B->own;
    int st = wig_order_streams(B->stream, next);
    if (st < 0) return st;
    B->stream = next;
    return WIG_OK;
}

int wig_batch_step(WigBatch *b, const WigParams *params, const WigStepCtl *ctl) {
    if (!b || !params || !ctl) return fail(WIG_E_INVALID, "null argument");
    return batch_step(*(Batch *)b, params, ctl);
}

int wig_batch_report(WigBatch *b, uint32_t *stats) {
    if (!b || !stats) return fail(WIG_E_INVALID, "null argument");
    return batch_report(*(Batch *)b, stats);
}

int wig_batch_report_async(WigBatch *b, uint64_t *ticket) {
    if (!b) return fail(WIG_E_INVALID, "null batch");
    return batch_report_async(*(Batch *)b, ticket);
}

int wig_batch_report_take(WigBatch *b, uint64_t ticket, uint32_t *stats, int wait) {
    if (!b) return fail(WIG_E_INVALID, "null batch");
    return batch_report_take(*(Batch *)b, ticket, stats, wait);
}

int wig_batch_sync(WigBatch *b) {
    if (!b) return fail(WIG_E_INVALID, "null batch");
    HIPCHECK(hipStreamSynchronize(((Batch *)b)->stream));
    return WIG_OK;
}

int wig_set_graphs(WigWorld *w, int on) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    w->solo->graphs = on != 0;
    return WIG_OK;
}

int wig_batch_set_graphs(WigBatch *b, int on) {
    if (!b) return fail(WIG_E_INVALID, "null batch");
    ((Batch *)b)->graphs = on != 0;
    return WIG_OK;
}

int wig_capture(WigWorld *w, int on) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    Dev &d = w->d;
    if (on && !w->check64) {
        const uint64_t P = 64ull * d.B;   // block-major, as the packed fine level
        uint64_t bytes = 4ull * P + 32ull * P + 32ull * d.B + 256u;
        if (hipMalloc(&w->check64, bytes) != hipSuccess) return fail(WIG_E_NOMEM, "capture arrays: %llu bytes", (unsigned long long)bytes);
        if (zero(stream_of(w), w->check64, bytes) != WIG_OK) return WIG_E_DEVICE;
        uint8_t *c = (uint8_t *)w->check64;
        d.head64 = (double *)c;
        d.part64 = (double *)(c + 256u);
        d.x64 = (double *)(c + 256u + 32ull * d.B);
        d.r64 = d.x64 + P;
        d.p64 = d.r64 + P;
        d.q64 = d.p64 + P;
        d.capb = (float *)(d.q64 + P);
    }
    if (!on) d.capb = nullptr;
    else d.capb = (float *)(d.q64 + 64ull * d.B);
    return WIG_OK;
}

int wig_cg64(WigWorld *w, double tol, uint32_t maxit, WigCheck64 *out) {
    if (!w || !out) return fail(WIG_E_INVALID, "null argument");
    if (!w->d.capb) return fail(WIG_E_INVALID, "wig_cg64 needs a solve run with wig_capture on");
    Batch &b = *w->solo;
    int st = batch_maps(b);
    if (st < 0) return st;
    b.htab[0] = w->d;
    hipStream_t s = b.stream;
    HIPCHECK(hipMemcpyAsync(b.tab, b.htab.data(), sizeof(Dev), hipMemcpyHostToDevice, s));
    WorldGrid g4 = grid_of(b, SH_BG4), g1 = grid_of(b, SH_ONE);
    dim3 n4(b.total[SH_BG4]);
    hipLaunchKernelGGL(k64_start, n4, dim3(256), 0, s, g4);
    hipLaunchKernelGGL(k64_sum, dim3(1), dim3(1024), 0, s, g1, 0);
    double h[16];
    uint32_t it = 0;
    for (; it < maxit; it++) {
        if (it % 32u == 0u) {
            HIPCHECK(hipMemcpyAsync(h, w->d.head64, 16, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            if (h[0] <= tol * tol * h[1]) break;
        }
        hipLaunchKernelGGL(k64_apply, n4, dim3(256), 0, s, g4);
        hipLaunchKernelGGL(k64_sum, dim3(1), dim3(1024), 0, s, g1, 1);
        hipLaunchKernelGGL(k64_update, n4, dim3(256), 0, s, g4);
        hipLaunchKernelGGL(k64_sum, dim3(1), dim3(1024), 0, s, g1, 2);
        hipLaunchKernelGGL(k64_direction, n4, dim3(256), 0, s, g4);
    }
    hipLaunchKernelGGL(k64_check, n4, dim3(256), 0, s, g4, 0);
    hipLaunchKernelGGL(k64_check, n4, dim3(256), 0, s, g4, 1);
    hipLaunchKernelGGL(k64_keep, dim3(1), dim3(1024), 0, s, g1, 2, 5);
    hipLaunchKernelGGL(k64_residual, n4, dim3(256), 0, s, g4, 0);
    hipLaunchKernelGGL(k64_keep, dim3(1), dim3(1024), 0, s, g1, 1, 7);
    hipLaunchKernelGGL(k64_residual, n4, dim3(256), 0, s, g4, 1);
    hipLaunchKernelGGL(k64_keep, dim3(1), dim3(1024), 0, s, g1, 1, 8);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(h, w->d.head64, sizeof h, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    out->iterations = it;
    out->pad = 0;
    double bb = h[1] > 0.0 ? h[1] : 1.0;
    out->relres64 = std::sqrt(h[8] / bb);
    out->err = h[6] > 0.0 ? std::sqrt(h[5] / h[6]) : 0.0;
    out->relres = std::sqrt(h[7] / bb);
    return WIG_OK;
}

int wig_cg64_x(WigWorld *w, double *x) {
    if (!w || !x) return fail(WIG_E_INVALID, "null argument");
    if (!w->check64) return fail(WIG_E_INVALID, "no FP64 solution (wig_cg64)");
    const Dev &d = w->d;
    std::vector<double> blocked(64ull * d.B);
    HIPCHECK(hipMemcpyAsync(blocked.data(), d.x64, 8ull * blocked.size(), hipMemcpyDeviceToHost, stream_of(w)));
    HIPCHECK(hipStreamSynchronize(stream_of(w)));
    // block-major on the device, row-major for the caller
    for (uint32_t y = 0; y < (uint32_t)d.H; y++)
        for (uint32_t x0 = 0; x0 < (uint32_t)d.W; x0++)
            x[(uint64_t)y * d.W + x0] = blocked[((uint64_t)(y >> 3) * d.bw + (x0 >> 3)) * 64u + ((y & 7u) << 3) + (x0 & 7u)];
    return WIG_OK;
}

int wig_batch_time_kernel(WigBatch *b, int id) {
    if (!b) return fail(WIG_E_INVALID, "null batch");
    return batch_time_kernel(*(Batch *)b, id);
}

int wig_batch_kernel_time(WigBatch *b, double *ms, uint32_t *launches) {
    if (!b) return fail(WIG_E_INVALID, "null batch");
    return batch_kernel_time(*(Batch *)b, ms, launches);
}

}  // extern "C"
