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

#include "../../include/wig.h"
#include "wig_device.h"

// the kernels (wig_kernels.hip)
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v);
__global__ void k_setStep(StepParams *sp, uint32_t frame, uint32_t seed, float earthG);
__global__ void k_sync(Dev d, int claims);
__global__ void k_syncParticles(Dev d);
__global__ void k_syncSeed(Dev d);
__global__ void k_compact(Dev d);
__global__ void k_retire(Dev d);
__global__ void k_clear(Dev d);
__global__ void k_gather(Dev d);
__global__ void k_localJoin(Dev d, int mode);
__global__ void k_edgeJoin(Dev d, int mode);
__global__ void k_sealAnchor(Dev d);
__global__ void k_classify(Dev d);
__global__ void k_setFaces(Dev d);
__global__ void k_divergence(Dev d);
__global__ void k_relax(Dev d, uint32_t cycle, int shift, int post);
__global__ void k_restrict(Dev d, uint32_t cycle);
__global__ void k_levelRho(Dev d);
__global__ void k_levelDown(Dev d, uint32_t cycle, uint32_t l, uint32_t T);
__global__ void k_coarseLDS(Dev d, uint32_t cycle);
__global__ void k_levelUp(Dev d, uint32_t cycle, uint32_t l, uint32_t T);
__global__ void k_cgStart(Dev d);
__global__ void k_cgAnchor(Dev d);
__global__ void k_cgResidual(Dev d);
__global__ void k_sumNorm(Dev d);
__global__ void k_cgPrepare(Dev d);
__global__ void k_cgRho(Dev d, uint32_t cycle);
__global__ void k_cgBeta(Dev d, uint32_t cycle);
__global__ void k_cgDirection(Dev d, uint32_t cycle);
__global__ void k_cgApply(Dev d, uint32_t cycle);
__global__ void k_cgAlpha(Dev d, uint32_t cycle);
__global__ void k_cgUpdate(Dev d, uint32_t cycle);
__global__ void k_cgCheck(Dev d, uint32_t cycle);
__global__ void k_cgFinish(Dev d, int ran);
__global__ void k_project(Dev d);
__global__ void k_extrapolate(Dev d, int layer);
__global__ void k_advect(Dev d);
__global__ void k_sortCount(Dev d);
__global__ void k_scanBlocks(Dev d);
__global__ void k_scanTotals(Dev d);
__global__ void k_scanFinish(Dev d);
__global__ void k_scatter(Dev d);
__global__ void k_rasterClaims(Dev d);
__global__ void k_raster(Dev d);
__global__ void k_unpaint(Dev d);

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
#define HIPCHECK(x)                                                                         \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return fail(WIG_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// Kernel ids for wig_time_kernel, in launch order.
#define WIG_KERNELS(X) X(setStep) X(sync) X(syncParticles) X(syncSeed) X(compact) X(retire) X(clear) X(gather) \
    X(localJoin) X(edgeJoin) X(sealAnchor) X(classify) X(setFaces) X(divergence) X(cgStart) X(cgAnchor) \
    X(cgResidual) X(sumNorm) X(cgPrepare) X(levelRho) X(relax) X(restrict) X(levelDown) X(coarseLDS) X(levelUp) \
    X(cgRho) X(cgBeta) X(cgDirection) X(cgApply) X(cgAlpha) X(cgUpdate) X(cgCheck) \
    X(cgFinish) X(project) X(extrapolate) X(advect) X(sortCount) X(scanBlocks) X(scanTotals) X(scanFinish) \
    X(scatter) X(rasterClaims) X(raster) X(unpaint)
#define X_ENUM(n) KID_##n,
enum { WIG_KERNELS(X_ENUM) KID_COUNT };
#define X_NAME(n) #n,
static const char *g_kernel_names[] = {WIG_KERNELS(X_NAME)};
#define MAX_TIMED 8192

struct WigWorld {
    Dev d;                       // sizes and device pointers
    uint64_t bytes[WIG_PART_COUNT];
    void *part[WIG_PART_COUNT];
    uint8_t *kind;               // 256 element kinds
    uint32_t *scratch;           // B words
    void *fine;                  // the solve's packed fine level (Dev::coef ... r2)
    unsigned long long *totals;  // WigRunTotals on the device
    StepParams *sp;              // the step's Params words
    size_t lds_bytes;            // k_coarseLDS's dynamic LDS
    uint32_t tile[9];            // k_levelDown/k_levelUp's tile side per level
    uint32_t *flag_host;         // pinned: the solve's iterations and convergence flag, read between chunks
    uint32_t last_cycles;        // iterations the last chunked solve ran
    hipStream_t own, stream;
    int device;
    int timed;                   // kernel id timed, -1 none
    uint32_t ntimed;
    hipEvent_t ev[2 * MAX_TIMED];
    bool ev_made;
};

// A launch, bracketed by events when its kernel is the one being timed.
template <typename... KArgs, typename... Args>
static void launch_lds(WigWorld *w, int id, void (*k)(KArgs...), dim3 grid, dim3 block, size_t lds, Args... args) {
    bool t = w->timed == id && w->ntimed < MAX_TIMED;
    if (t) hipEventRecord(w->ev[2 * w->ntimed], w->stream);
    hipLaunchKernelGGL(k, grid, block, lds, w->stream, args...);
    if (t) hipEventRecord(w->ev[2 * w->ntimed + 1], w->stream), w->ntimed++;
}
template <typename... KArgs, typename... Args>
static void launch(WigWorld *w, int id, void (*k)(KArgs...), dim3 grid, dim3 block, Args... args) {
    launch_lds(w, id, k, grid, block, 0, args...);
}

static uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }

// k_levelDown/k_levelUp: tiles of T x T level cells, each with a halo of
// one cell a half-sweep (k_levelDown one more), in LDS 7 floats a cell, and
// k_levelDown's residuals (2 a tile cell) or k_levelUp's coarse cells (3)
#define MAX_LEVEL_PAIRS 8u
static dim3 tileGrid(const Dev &d, uint32_t l, uint32_t T) { return dim3(ceil_div(d.lw[l], T), ceil_div(d.lh[l], T)); }
static size_t downLds(const Dev &d, uint32_t T) {
    size_t R = T + 2u * (2u * d.level_pairs + 1u);
    return 4u * (7u * R * R + 2u * T * T);
}
static size_t upLds(const Dev &d, uint32_t T) {
    size_t R = T + 4u * d.level_pairs, Rc = R / 2u + 3u;
    return 4u * (7u * R * R + 3u * Rc * Rc);
}
// the world's tile for level l, narrowed to fit 64 KB of LDS
static uint32_t levelTileSide(const WigWorld *w, const Dev &d, uint32_t l) {
    uint32_t T = w->tile[l];
    while (T > 8u && std::max(downLds(d, T), upLds(d, T)) > 65536u) T /= 2u;
    return T;
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
    w->bytes[WIG_PART_SOLVE] = sizeof(SolveHead) + 32ull * d.B;
}

extern "C" {

const char *wig_last_error(void) { return g_err.c_str(); }

int wig_device_ok(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return 0;
    return 1;
}

static int zero(WigWorld *w, void *p, uint64_t bytes) {
    uint64_t words = bytes / 4u;
    if (!words) return WIG_OK;
    uint32_t grid = (uint32_t)std::min<uint64_t>((words + 255u) / 256u, 65535u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, w->stream, (uint32_t *)p, words, 0u);
    HIPCHECK(hipGetLastError());
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
    memset(&w->d, 0, sizeof w->d);
    size_world(w, width, height, coarse_levels);
    if (w->d.lw[w->d.L] * w->d.lh[w->d.L] > (uint32_t)MG_COARSEST) {
        fail(WIG_E_INVALID, "the coarsest level (%ux%u) holds more than %d cells", w->d.lw[w->d.L], w->d.lh[w->d.L], MG_COARSEST);
        delete w;
        return nullptr;
    }
    if (hipGetDevice(&w->device) != hipSuccess || hipStreamCreate(&w->own) != hipSuccess) {
        fail(WIG_E_DEVICE, "no HIP device");
        delete w;
        return nullptr;
    }
    w->stream = w->own;
    for (int p = 0; p < WIG_PART_COUNT; p++) {
        if (hipMalloc(&w->part[p], w->bytes[p]) != hipSuccess) {
            fail(WIG_E_NOMEM, "hipMalloc of part %d (%llu bytes) failed", p, (unsigned long long)w->bytes[p]);
            wig_destroy(w);
            return nullptr;
        }
    }
    w->timed = -1;
    if (hipMalloc((void **)&w->kind, 256) != hipSuccess || hipMalloc((void **)&w->scratch, 4ull * w->d.B) != hipSuccess
        || hipMalloc(&w->fine, 41ull * w->d.C + 256u) != hipSuccess
        || hipMalloc((void **)&w->totals, sizeof(WigRunTotals)) != hipSuccess
        || hipMalloc((void **)&w->sp, sizeof(StepParams)) != hipSuccess
        || hipHostMalloc((void **)&w->flag_host, 64, hipHostMallocDefault) != hipSuccess) {
        fail(WIG_E_NOMEM, "hipMalloc failed");
        wig_destroy(w);
        return nullptr;
    }
    for (int p = 0; p < WIG_PART_COUNT; p++)
        if (zero(w, w->part[p], w->bytes[p]) != WIG_OK) { wig_destroy(w); return nullptr; }
    zero(w, w->kind, 256);
    zero(w, w->fine, 41ull * w->d.C + 256u);
    zero(w, w->totals, sizeof(WigRunTotals));
    zero(w, w->sp, sizeof(StepParams));
    Dev &d = w->d;
    d.totals = w->totals;
    d.sp = w->sp;
    // the coarse levels one workgroup's LDS holds (k_coarseLDS): from the
    // first that fits with all below it
    int lds_max = 65536;
    hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, w->device);
    d.lds_level = d.L;
    for (uint32_t l = d.L; l >= 1; l--) {
        size_t bytes = 0;
        for (uint32_t k = l; k <= d.L; k++) bytes += 4ull * LDS_LEVEL_FLOATS * d.lw[k] * d.lh[k];
        if (bytes > (size_t)lds_max) break;
        d.lds_level = l;
        w->lds_bytes = bytes;
    }
    if (w->lds_bytes == 0) w->lds_bytes = 4ull * LDS_LEVEL_FLOATS * d.lw[d.L] * d.lh[d.L];
    for (uint32_t l = 0; l < 9; l++) w->tile[l] = 16u;
    if (w->lds_bytes > 65536 && hipFuncSetAttribute((const void *)k_coarseLDS, hipFuncAttributeMaxDynamicSharedMemorySize, (int)w->lds_bytes) != hipSuccess) {
        fail(WIG_E_DEVICE, "k_coarseLDS cannot take %zu bytes of LDS", w->lds_bytes);
        wig_destroy(w);
        return nullptr;
    }
    d.cells = (uint32_t *)w->part[WIG_PART_CELLS];
    d.sorted = (uint32_t *)w->part[WIG_PART_PARTICLES0];
    d.parts = (uint32_t *)w->part[WIG_PART_PARTICLES1];
    d.counters = (uint32_t *)w->part[WIG_PART_COUNTERS];
    d.field = (float *)w->part[WIG_PART_FIELD];
    d.blocks = (uint32_t *)w->part[WIG_PART_BLOCKS];
    d.cg = (float *)w->part[WIG_PART_FLUIDCG];
    d.head = (SolveHead *)w->part[WIG_PART_SOLVE];
    d.bpart = (double *)((uint8_t *)w->part[WIG_PART_SOLVE] + sizeof(SolveHead));
    d.kind = w->kind;
    d.scratch = w->scratch;
    {
        uint8_t *f = (uint8_t *)w->fine;
        d.coef = (float4 *)f;
        d.pa = (float2 *)(f + 16ull * d.C);
        d.pb = (float2 *)(f + 24ull * d.C);
        d.r2 = (float2 *)(f + 32ull * d.C);
        d.cflag = f + 40ull * d.C;
    }
    if (hipStreamSynchronize(w->stream) != hipSuccess) {
        fail(WIG_E_DEVICE, "stream failed at creation");
        wig_destroy(w);
        return nullptr;
    }
    return w;
}

int wig_destroy(WigWorld *w) {
    if (!w) return WIG_OK;
    if (w->stream) hipStreamSynchronize(w->stream);
    for (int p = 0; p < WIG_PART_COUNT; p++)
        if (w->part[p]) hipFree(w->part[p]);
    if (w->kind) hipFree(w->kind);
    if (w->scratch) hipFree(w->scratch);
    if (w->fine) hipFree(w->fine);
    if (w->totals) hipFree(w->totals);
    if (w->sp) hipFree(w->sp);
    if (w->flag_host) hipHostFree(w->flag_host);
    if (w->ev_made)
        for (int i = 0; i < 2 * MAX_TIMED; i++) hipEventDestroy(w->ev[i]);
    if (w->own) hipStreamDestroy(w->own);
    delete w;
    return WIG_OK;
}

int wig_set_kinds(WigWorld *w, const uint8_t kinds[256]) {
    if (!w || !kinds) return fail(WIG_E_INVALID, "null argument");
    HIPCHECK(hipMemcpyAsync(w->kind, kinds, 256, hipMemcpyHostToDevice, w->stream));
    HIPCHECK(hipStreamSynchronize(w->stream));
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

int wig_upload(WigWorld *w, int part, const void *src, size_t bytes) {
    if (!w || !src || part < 0 || part >= WIG_PART_COUNT) return fail(WIG_E_INVALID, "bad part %d", part);
    if (bytes != w->bytes[part]) return fail(WIG_E_INVALID, "part %d is %llu bytes, not %zu", part, (unsigned long long)w->bytes[part], bytes);
    HIPCHECK(hipMemcpyAsync(w->part[part], src, bytes, hipMemcpyHostToDevice, w->stream));
    HIPCHECK(hipStreamSynchronize(w->stream));
    return WIG_OK;
}

int wig_download(WigWorld *w, int part, void *dst, size_t bytes) {
    if (!w || !dst || part < 0 || part >= WIG_PART_COUNT) return fail(WIG_E_INVALID, "bad part %d", part);
    if (bytes != w->bytes[part]) return fail(WIG_E_INVALID, "part %d is %llu bytes, not %zu", part, (unsigned long long)w->bytes[part], bytes);
    HIPCHECK(hipMemcpyAsync(dst, w->part[part], bytes, hipMemcpyDeviceToHost, w->stream));
    HIPCHECK(hipStreamSynchronize(w->stream));
    return WIG_OK;
}

int wig_device_ptr(WigWorld *w, int part, void **ptr) {
    if (!w || !ptr || part < 0 || part >= WIG_PART_COUNT) return fail(WIG_E_INVALID, "bad part %d", part);
    *ptr = w->part[part];
    return WIG_OK;
}

int wig_set_stream(WigWorld *w, void *stream) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    w->stream = stream ? (hipStream_t)stream : w->own;
    return WIG_OK;
}

int wig_sync(WigWorld *w) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    HIPCHECK(hipStreamSynchronize(w->stream));
    return WIG_OK;
}

int wig_report(WigWorld *w, uint32_t stats[WIG_STATS]) {
    if (!w || !stats) return fail(WIG_E_INVALID, "null argument");
    uint32_t *src = (uint32_t *)w->part[WIG_PART_BLOCKS] + 4ull * w->d.B;
    HIPCHECK(hipMemcpyAsync(stats, src, 4u * WIG_STATS, hipMemcpyDeviceToHost, w->stream));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, w->stream, src + WIG_STAT_MOST_CYCLES, (uint64_t)1, 0u);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(w->stream));
    return WIG_OK;
}

int wig_run_totals(WigWorld *w, WigRunTotals *out) {
    if (!w || !out) return fail(WIG_E_INVALID, "null argument");
    HIPCHECK(hipMemcpyAsync(out, w->totals, sizeof *out, hipMemcpyDeviceToHost, w->stream));
    HIPCHECK(hipStreamSynchronize(w->stream));
    return WIG_OK;
}

int wig_time_kernel(WigWorld *w, int id) {
    if (!w || id > KID_COUNT) return fail(WIG_E_INVALID, "bad kernel id %d", id);
    if (!w->ev_made) {
        for (int i = 0; i < 2 * MAX_TIMED; i++) HIPCHECK(hipEventCreate(&w->ev[i]));
        w->ev_made = true;
    }
    w->timed = id < 0 ? -1 : id;
    w->ntimed = 0;
    return WIG_OK;
}

int wig_kernel_time(WigWorld *w, double *ms, uint32_t *launches) {
    if (!w || !ms || !launches) return fail(WIG_E_INVALID, "null argument");
    HIPCHECK(hipStreamSynchronize(w->stream));
    double total = 0.0;
    for (uint32_t i = 0; i < w->ntimed; i++) {
        float t = 0.0f;
        HIPCHECK(hipEventElapsedTime(&t, w->ev[2 * i], w->ev[2 * i + 1]));
        total += t;
    }
    *ms = total;
    *launches = w->ntimed;
    w->ntimed = 0;
    return WIG_OK;
}

int wig_set_tile(WigWorld *w, uint32_t level, uint32_t side) {
    if (!w) return fail(WIG_E_INVALID, "null world");
    if (level < 1u || level >= w->d.lds_level) return fail(WIG_E_INVALID, "level %u is not swept by tiles (1 .. %u)", level, w->d.lds_level - 1u);
    if (side != 8u && side != 16u && side != 32u) return fail(WIG_E_INVALID, "tile side %u: 8, 16 or 32", side);
    w->tile[level] = side;
    return WIG_OK;
}

const char *wig_kernel_name(int id) { return id >= 0 && id < KID_COUNT ? g_kernel_names[id] : (id == KID_COUNT ? "solve" : ""); }
int wig_kernel_count(void) { return KID_COUNT + 1; }

int wig_solve_stats(WigWorld *w, WigSolveStats *out) {
    if (!w || !out) return fail(WIG_E_INVALID, "null argument");
    HIPCHECK(hipMemcpyAsync(out, w->part[WIG_PART_SOLVE], sizeof *out, hipMemcpyDeviceToHost, w->stream));
    HIPCHECK(hipStreamSynchronize(w->stream));
    return WIG_OK;
}

// Grid of the passes over the blocks worked on: they stride over the active
// list, so a few workgroups a CU cover any count.
static uint32_t block_grid(const Dev &d) { return std::min<uint32_t>(d.B + d.bw + d.bh + 1u, 4096u); }
static uint32_t particle_grid(const Dev &d) { return std::min<uint32_t>((d.cap + 255u) / 256u, 4096u); }

int wig_fluid_step(WigWorld *w, const WigParams *P, const WigStepCtl *ctl) {
    if (!w || !P || !ctl) return fail(WIG_E_INVALID, "null argument");
    Dev d = w->d;
    if (P->size[0] != (uint32_t)d.W || P->size[1] != (uint32_t)d.H)
        return fail(WIG_E_INVALID, "params size %ux%u is not the world's %dx%d", P->size[0], P->size[1], d.W, d.H);
    if (ctl->viscous || ctl->surface || ctl->bodies_run)
        return fail(WIG_E_INVALID, "viscosity, surface tension and bodies are not built (water only)");
    if (ctl->fluid_cycles > WIG_MAX_CYCLES || ctl->sweep_pairs > 64u || ctl->coarse_sweeps > 4096u || ctl->tol_mode > 1u
        || ctl->level_pairs > MAX_LEVEL_PAIRS)
        return fail(WIG_E_INVALID, "solver control out of range");
    if (ctl->pre_relax != 1u || ctl->post_relax != 1u)
        return fail(WIG_E_INVALID, "one relaxation before and one after the coarse levels (PRE_RELAX = POST_RELAX = 1) is what is built");
    d.sweep_pairs = ctl->sweep_pairs;
    d.coarse_sweeps = ctl->coarse_sweeps;
    d.tol_mode = ctl->tol_mode;
    d.tol = ctl->tol;
    d.level_pairs = ctl->level_pairs ? ctl->level_pairs : 1u;
    hipStream_t s = w->stream;
    dim3 world(d.bw, d.bh), b64(64), b256(256);
    uint32_t bg = block_grid(d), pg = particle_grid(d), bg4 = std::min<uint32_t>((d.B + 3u) / 4u, 1024u);
    launch(w, KID_setStep, k_setStep, dim3(1), dim3(1), w->sp, P->frame, P->seed, P->earthG);
    if (!ctl->full) {
        launch(w, KID_sync, k_sync, world, b64, d, 0);
        launch(w, KID_compact, k_compact, dim3(1), dim3(1024), d);
        launch(w, KID_retire, k_retire, dim3(bg), b64, d);
        HIPCHECK(hipGetLastError());
        return WIG_OK;
    }
    launch(w, KID_sync, k_sync, world, b64, d, 1);
    launch(w, KID_syncParticles, k_syncParticles, dim3(pg), b256, d);
    launch(w, KID_syncSeed, k_syncSeed, world, b64, d);
    launch(w, KID_compact, k_compact, dim3(1), dim3(1024), d);
    launch(w, KID_retire, k_retire, dim3(bg), b64, d);
    launch(w, KID_clear, k_clear, dim3(bg), b64, d);
    launch(w, KID_gather, k_gather, dim3(pg), b256, d);
    launch(w, KID_localJoin, k_localJoin, dim3(bg), b64, d, 0);
    launch(w, KID_edgeJoin, k_edgeJoin, dim3(bg), b64, d, 0);
    launch(w, KID_sealAnchor, k_sealAnchor, dim3(bg), b64, d);
    launch(w, KID_classify, k_classify, dim3(bg), b64, d);
    launch(w, KID_setFaces, k_setFaces, dim3(bg), b64, d);
    launch(w, KID_divergence, k_divergence, dim3(bg), b64, d);
    // (the solve, cgStart .. cgFinish, timed as one: the pseudo-kernel "solve")
    bool timeSolve = w->timed == KID_COUNT && w->ntimed < MAX_TIMED;
    if (timeSolve) hipEventRecord(w->ev[2 * w->ntimed], s);
    launch(w, KID_cgStart, k_cgStart, dim3(bg), b64, d);
    launch(w, KID_localJoin, k_localJoin, dim3(bg), b64, d, 1);
    launch(w, KID_edgeJoin, k_edgeJoin, dim3(bg), b64, d, 1);
    launch(w, KID_cgAnchor, k_cgAnchor, dim3(bg), b64, d);
    launch(w, KID_cgResidual, k_cgResidual, dim3(bg4), b256, d);
    launch(w, KID_sumNorm, k_sumNorm, dim3(1), dim3(1024), d);
    launch(w, KID_cgPrepare, k_cgPrepare, dim3(bg), b64, d);
    dim3 lg1(ceil_div(d.lw[1], 8u), ceil_div(d.lh[1], 8u));
    if (ctl->fluid_cycles > 0u) launch(w, KID_levelRho, k_levelRho, lg1, b64, d);
    bool shifted = ctl->shifted_first != 0;
    uint32_t l0 = d.lds_level;
    uint32_t first = ctl->first_chunk == WIG_CHUNK_AUTO ? std::max<uint32_t>(w->last_cycles, 1u) : ctl->first_chunk;
    uint32_t next_check = first ? first : WIG_MAX_CYCLES + 1u;
    for (uint32_t cycle = 0; cycle < ctl->fluid_cycles; cycle++) {
        if (cycle == next_check) {
            // the chunk encoded so far has run: stop if the solve converged
            HIPCHECK(hipMemcpyAsync(w->flag_host, &w->d.head->cycles, 8, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            if (w->flag_host[1]) {
                w->last_cycles = w->flag_host[0];
                break;
            }
            next_check = cycle + (ctl->chunk ? ctl->chunk : 1u);
        }
        launch(w, KID_relax, k_relax, dim3(bg), b64, d, cycle, shifted ? WIG_BLOCK / 2 : 0, 0);
        shifted = !shifted;
        launch(w, KID_restrict, k_restrict, lg1, b64, d, cycle);
        for (uint32_t l = 1; l < l0; l++) {   // level l's sweeps, and restrictLevel into l + 1
            uint32_t T = levelTileSide(w, d, l);
            launch_lds(w, KID_levelDown, k_levelDown, tileGrid(d, l, T), dim3(256), downLds(d, T), d, cycle, l, T);
        }
        launch_lds(w, KID_coarseLDS, k_coarseLDS, dim3(1), dim3(1024), w->lds_bytes, d, cycle);
        for (uint32_t l = l0 - 1; l >= 1; l--) {
            uint32_t T = levelTileSide(w, d, l);
            launch_lds(w, KID_levelUp, k_levelUp, tileGrid(d, l, T), dim3(256), upLds(d, T), d, cycle, l, T);
        }
        launch(w, KID_relax, k_relax, dim3(bg), b64, d, cycle, shifted ? WIG_BLOCK / 2 : 0, 1);
        shifted = !shifted;
        launch(w, KID_cgRho, k_cgRho, dim3(bg4), b256, d, cycle);
        launch(w, KID_cgBeta, k_cgBeta, dim3(1), dim3(1024), d, cycle);
        launch(w, KID_cgDirection, k_cgDirection, dim3(bg), b64, d, cycle);
        launch(w, KID_cgApply, k_cgApply, dim3(bg4), b256, d, cycle);
        launch(w, KID_cgAlpha, k_cgAlpha, dim3(1), dim3(1024), d, cycle);
        launch(w, KID_cgUpdate, k_cgUpdate, dim3(bg4), b256, d, cycle);
        launch(w, KID_cgCheck, k_cgCheck, dim3(1), dim3(1024), d, cycle);
    }
    launch(w, KID_cgFinish, k_cgFinish, dim3(bg), b64, d, ctl->fluid_cycles > 0u ? 1 : 0);
    if (timeSolve) hipEventRecord(w->ev[2 * w->ntimed + 1], s), w->ntimed++;
    launch(w, KID_project, k_project, dim3(bg), b64, d);
    launch(w, KID_extrapolate, k_extrapolate, dim3(bg), b64, d, 1);
    launch(w, KID_extrapolate, k_extrapolate, dim3(bg), b64, d, 2);
    launch(w, KID_advect, k_advect, dim3(pg), b256, d);
    launch(w, KID_sortCount, k_sortCount, dim3(pg), b256, d);
    launch(w, KID_scanBlocks, k_scanBlocks, dim3(bg), b64, d);
    launch(w, KID_scanTotals, k_scanTotals, dim3(1), b256, d);
    launch(w, KID_scanFinish, k_scanFinish, dim3(bg), b64, d);
    launch(w, KID_scatter, k_scatter, dim3(pg), b256, d);
    launch(w, KID_rasterClaims, k_rasterClaims, dim3(bg), b64, d);
    launch(w, KID_raster, k_raster, dim3(bg), b64, d);
    launch(w, KID_unpaint, k_unpaint, dim3(bg), b64, d);
    HIPCHECK(hipGetLastError());
    return WIG_OK;
}

}  // extern "C"
