// The air's projection on gfx950: the MG-PCG that solves for the air's
// dynamic pressure (shaders/air.wgsl:869-1536), in the order
// Simulation.encodeAir dispatches it (js/sim.js:2264-2289), on the
// reference's buffers: airProjection (mg: every level, two float4s a
// cell), airCg (cg: a float4 a level-0 cell, then the sums), airFastest
// (status words).
//
// The V-cycle's level passes sweep tiles in LDS: a T x T square of a
// level's cells with HALO = 5 more each side (two red-black pairs leave
// each sweep one more ring in from the edge stale, and the restriction
// reads one ring beyond the square), so that the square comes out as whole
// sweeps of the level would leave it.  The sums are the reference's fixed
// trees (a group of 64 level-0 cells, halving; 256 strided lanes, halving),
// their last steps as wave shuffles, so every run gives the same bits.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>

#include "../../include/wig.h"

#define AIR_LEVELS 9        // js/sim.js:28
#define AIR_COARSEST 256    // js/shared.js:239
#define ANCHOR 1e-6f        // air.wgsl:146
#define COARSE_SWEEPS 20    // air.wgsl:153
#define COARSE_OMEGA 1.7f   // air.wgsl:154
#define COARSEN 1.8f        // air.wgsl:889
#define CORRECT 1.5f        // air.wgsl:901
#define AIR_TOLERANCE 1e-4f // air.wgsl:945
#define AIR_SMOOTH 2        // air.wgsl:1214
#define HALO (2 * AIR_SMOOTH + 1)

struct AirDev {
    float4 *mg;
    float4 *cg;
    uint32_t *status;
    uint32_t *most;     // a level-0 pass's most |r| bits, a workgroup each (this library's)
    float *pre0;        // level 0's smoothed correction, down to up (this library's; none where held)
    float *pnew;        // cgDirection's p, cgApply to cgUpdate (this library's)
    // Level 0 during a solve (this library's): its record's words the solve
    // does not change (right and bottom weights, edge, fixed), its rhs (-r)
    // and the V-cycle's answer z.  The records' potential and rhs are
    // written once, at the end (k_airFinish), but on a grid of one level,
    // whose coarsest solve reads them every iteration.
    float4 *w0;
    float *rhs0, *z0;
    const float *mean;
    uint32_t L, W, H, gx, groups;
    uint32_t w[AIR_LEVELS], h[AIR_LEVELS];
    uint32_t base[AIR_LEVELS];   // levelBase, in float4s
};

__device__ __forceinline__ uint32_t mgAt(const AirDev &a, uint32_t l, int32_t x, int32_t y) {
    return a.base[l] + 2u * ((uint32_t)y * a.w[l] + (uint32_t)x);
}
__device__ __forceinline__ float *sums(const AirDev &a) { return (float *)(a.cg + (uint64_t)a.W * a.H); }
__device__ __forceinline__ bool airConverged(const AirDev &a, uint32_t cycle) {   // :963-965
    return __uint_as_float(a.status[1u + (cycle + 2u) % 3u]) < AIR_TOLERANCE;
}
__device__ __forceinline__ float anchorAt(const AirDev &a, uint32_t l, int32_t cx, int32_t cy) {   // :912-916
    const int32_t lox = cx << l, loy = cy << l;
    const int32_t hx = min(lox + (1 << l), (int32_t)a.W), hy = min(loy + (1 << l), (int32_t)a.H);
    return ANCHOR * (float)((hx - lox) * (hy - loy)) / (l >= 1u ? COARSEN : 1.0f);
}
__device__ __forceinline__ uint32_t waveMaxU32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
// The most |r| of a workgroup of 256, into its word of `most`; k_airSlot
// then takes the most of those into the status slot (a max is the same in
// any order, and no workgroup waits on another's atomic)
__device__ __forceinline__ void groupMax(uint32_t *most, uint32_t v) {
    __shared__ uint32_t w4[4];
    v = waveMaxU32(v);
    if ((threadIdx.x & 63u) == 0u) w4[threadIdx.x / 64u] = v;
    __syncthreads();
    if (threadIdx.x == 0u) most[blockIdx.x] = max(max(w4[0], w4[1]), max(w4[2], w4[3]));
}
// the status slot's atomicMax of every cell's |r| (:1047, :1193): the most
// of the `groups` workgroups' words, into the slot as it stands
__global__ void __launch_bounds__(1024) k_airSlot(AirDev a, uint32_t slot, uint32_t groups, uint32_t cycle, int always) {
    if (!always && airConverged(a, cycle)) return;
    __shared__ uint32_t w16[16];
    uint32_t v = 0u;
    for (uint32_t j = threadIdx.x; j < groups; j += 1024u) v = max(v, a.most[j]);
    v = waveMaxU32(v);
    if ((threadIdx.x & 63u) == 0u) w16[threadIdx.x / 64u] = v;
    __syncthreads();
    if (threadIdx.x == 0u) {
        for (int w = 0; w < 16; w++) v = max(v, w16[w]);
        a.status[slot] = max(a.status[slot], v);
    }
}
// The level-0 passes over cells: a grid of at most CELL_GROUPS workgroups
// striding over them (so that few workgroups, not every wavefront, take
// the status slot's atomic)
#define CELL_GROUPS 1024u
#define UPDATE_HELD 4
// the 64-value tree (sumGroup, :1054-1069): lane i < s takes lane i + s's
__device__ __forceinline__ float groupTree(float v) {
    for (int s = 32; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
    return v;
}

// gather's clearing of the solve's slots (:444-449)
__global__ void k_airClear(AirDev a) {
    if (threadIdx.x == 0) {
        a.status[1] = 0u;
        a.status[3] = 0u;
        a.status[4] = 0u;
    }
}

// applyA (:969-993) on cg's component `which`
template <int WHICH>
__device__ __forceinline__ float applyA(const AirDev &a, int32_t x, int32_t y, float4 m0, float4 m1) {
    const uint32_t i = (uint32_t)y * a.W + (uint32_t)x;
    const float *c = (const float *)a.cg;
    float diag = m0.z + m0.w + m1.x + anchorAt(a, 0u, x, y);
    float sum = 0.0f;
    if ((uint32_t)x + 1u < a.W) sum += m0.z * c[4u * (i + 1u) + WHICH];
    if ((uint32_t)y + 1u < a.H) sum += m0.w * c[4u * (i + a.W) + WHICH];
    if (x > 0) {
        const float w = a.mg[mgAt(a, 0u, x - 1, y)].z;
        sum += w * c[4u * (i - 1u) + WHICH];
        diag += w;
    }
    if (y > 0) {
        const float w = a.mg[mgAt(a, 0u, x, y - 1)].w;
        sum += w * c[4u * (i - a.W) + WHICH];
        diag += w;
    }
    return diag * c[4u * i + WHICH] - sum;
}

// cgStart (:1027-1048)
__global__ void __launch_bounds__(256) k_airStart(AirDev a) {
    uint32_t most = 0u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < a.W * a.H; i += gridDim.x * 256u) {
        const int32_t x = (int32_t)(i % a.W), y = (int32_t)(i / a.W);
        const uint32_t at = mgAt(a, 0u, x, y);
        float4 m0 = a.mg[at];
        const float4 m1 = a.mg[at + 1u];
        float r = 0.0f;
        if (!(m1.y > 0.0f)) {
            if (m1.w > 0.0f) m0.y += a.mean[i];
            r = -m0.y - applyA<0>(a, x, y, m0, m1);
        }
        // (a pass writes no potential, so the neighbours' above are as they were)
        const float px = a.cg[i].x;
        a.cg[i] = make_float4(px, r, 0.0f, 0.0f);
        a.mg[at].x = 0.0f;
        a.mg[at].y = -r;
        a.w0[i] = make_float4(m0.z, m0.w, m1.x, m1.y);
        a.rhs0[i] = -r;
        a.z0[i] = 0.0f;
        most = max(most, __float_as_uint(fabsf(r)));
    }
    groupMax(a.most, most);
}

// The V-cycle's level pass on level l (down: smoothDown, :1428-1459; up:
// :1461-1474), a workgroup a T x T square of the level.
template <int T>
struct AirTile {
    static constexpr int SPAN = T + 2 * HALO, CELLS = SPAN * SPAN;
    float x[CELLS], rhs[CELLS], right[CELLS], down[CELLS], diag[CELLS];
};

template <int T, int NT, bool UP>
__device__ __forceinline__ void loadTile(const AirDev &a, AirTile<T> &t, uint32_t l, int32_t cx0, int32_t cy0) {   // :1255-1291
    constexpr int SPAN = AirTile<T>::SPAN, CELLS = AirTile<T>::CELLS, R = (CELLS + NT - 1) / NT;
    const int32_t dx = (int32_t)a.w[l], dy = (int32_t)a.h[l];
    // a thread's R cells: every load made first (clamped addresses, values
    // chosen after), so that they are in flight together
    float4 m0[R], m1[R];
    float pre[R], up[R];
    bool in[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int k = (int)threadIdx.x + r * NT;
        const int32_t x = cx0 + k % SPAN, y = cy0 + k / SPAN;
        in[r] = k < CELLS && x >= 0 && y >= 0 && x < dx && y < dy;
        const int32_t xc = in[r] ? x : 0, yc = in[r] ? y : 0;
        if (l == 0u) {
            const uint32_t i = (uint32_t)yc * a.W + (uint32_t)xc;
            const float4 w = a.w0[i];
            m0[r] = make_float4(0.0f, a.rhs0[i], w.x, w.y);
            m1[r] = make_float4(w.z, w.w, 0.0f, 0.0f);
        } else {
            const uint32_t at = mgAt(a, l, xc, yc);
            m0[r] = a.mg[at];
            m1[r] = a.mg[at + 1u];
        }
        if (UP) {
            pre[r] = l == 0u ? a.pre0[(uint32_t)yc * a.W + (uint32_t)xc] : 0.0f;
            up[r] = a.mg[mgAt(a, l + 1u, xc / 2, yc / 2)].x;
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int k = (int)threadIdx.x + r * NT;
        if (k >= CELLS) continue;
        const int32_t x = cx0 + k % SPAN, y = cy0 + k / SPAN;
        float vx = 0.0f, rhs = 0.0f, right = 0.0f, down = 0.0f, diag = -1.0f;
        if (in[r]) {
            rhs = m0[r].y;
            right = m0[r].z;
            down = m0[r].w;
            if (UP) vx = l == 0u ? pre[r] : m1[r].z;
            if (!(m1[r].y > 0.0f)) {
                if (UP) vx += CORRECT * up[r];
                diag = m0[r].z + m0[r].w + m1[r].x + anchorAt(a, l, x, y);
            }
        }
        t.x[k] = vx;
        t.rhs[k] = rhs;
        t.right[k] = right;
        t.down[k] = down;
        t.diag[k] = diag;
    }
    __syncthreads();
    // the left and upper faces' weights into the diagonal, from the tile
    // (the outermost ring's are never read: it is not swept)
    for (int k = (int)threadIdx.x; k < CELLS; k += NT) {
        const int tx = k % SPAN, ty = k / SPAN;
        if (tx == 0 || ty == 0 || !(t.diag[k] >= 0.0f)) continue;
        const int32_t x = cx0 + tx, y = cy0 + ty;
        float d = t.diag[k];
        if (x > 0) d += t.right[k - 1];
        if (y > 0) d += t.down[k - SPAN];
        t.diag[k] = d;
    }
    __syncthreads();
}

// tileSum (:1293-1309)
template <int T>
__device__ __forceinline__ float tileSum(const AirTile<T> &t, int k, int32_t x, int32_t y, int32_t dx, int32_t dy) {
    constexpr int SPAN = AirTile<T>::SPAN;
    float sum = 0.0f;
    if (x + 1 < dx) sum += t.right[k] * t.x[k + 1];
    if (y + 1 < dy) sum += t.down[k] * t.x[k + SPAN];
    if (x > 0) sum += t.right[k - 1] * t.x[k - 1];
    if (y > 0) sum += t.down[k - SPAN] * t.x[k - SPAN];
    return sum;
}

// sweepTile (:1313-1326): the cells of one colour but the outermost ring
template <int T, int NT>
__device__ __forceinline__ void sweepTile(AirTile<T> &t, int color, int32_t cx0, int32_t cy0, int32_t dx, int32_t dy) {
    constexpr int SPAN = AirTile<T>::SPAN, INNER = SPAN - 2;
    // the inner (SPAN - 2)^2 cells, half of them this colour: thread j takes
    // the j-th of the colour in row order
    for (int j = (int)threadIdx.x; j < (INNER * INNER + 1) / 2; j += NT) {
        const int q = 2 * j, row = q / INNER;
        int col = q % INNER;
        const int32_t y = cy0 + 1 + row;
        // the colour's cells of a row alternate; shift by one where the row's
        // first cell is the other colour
        if (((cx0 + 1 + y) & 1) != color) col += 1;
        if (col >= INNER) continue;
        const int32_t x = cx0 + 1 + col;
        const int k = (row + 1) * SPAN + col + 1;
        if (!(t.diag[k] >= 0.0f)) continue;
        t.x[k] = (tileSum(t, k, x, y, dx, dy) - t.rhs[k]) / t.diag[k];
    }
    __syncthreads();
}

// restrictCell (:1349-1417): level l + 1's cell (cx, cy) from its four,
// the fine cells' state from the tile
template <int T>
__device__ void restrictCell(const AirDev &a, const AirTile<T> &t, uint32_t l, int32_t cx0, int32_t cy0, int32_t cx, int32_t cy) {
    constexpr int SPAN = AirTile<T>::SPAN;
    const int32_t fdx = (int32_t)a.w[l], fdy = (int32_t)a.h[l];
    auto kOf = [&](int32_t x, int32_t y) { return (y - cy0) * SPAN + (x - cx0); };
    auto held = [&](int32_t x, int32_t y) { return t.diag[kOf(x, y)] < 0.0f; };
    float rhs = 0.0f, edge = 0.0f;
    uint32_t nfree = 0, nheld = 0;
    for (int q = 0; q < 4; q++) {
        const int32_t fx = 2 * cx + (q & 1), fy = 2 * cy + (q >> 1);
        if (fx >= fdx || fy >= fdy) continue;
        const int k = kOf(fx, fy);
        if (t.diag[k] < 0.0f) {
            nheld++;
            continue;
        }
        nfree++;
        edge += a.mg[mgAt(a, l, fx, fy) + 1u].x;
        if (fx + 1 < fdx && held(fx + 1, fy)) edge += t.right[k];
        if (fy + 1 < fdy && held(fx, fy + 1)) edge += t.down[k];
        if (fx > 0 && held(fx - 1, fy)) edge += t.right[k - 1];
        if (fy > 0 && held(fx, fy - 1)) edge += t.down[k - SPAN];
        rhs += t.rhs[k] - (tileSum(t, k, fx, fy, fdx, fdy) - t.diag[k] * t.x[k]);
    }
    const int32_t right = min(2 * cx + 1, fdx - 1), bottom = min(2 * cy + 1, fdy - 1);
    float wR = 0.0f, wD = 0.0f;
    for (int r = 0; r < 2; r++) {
        const int32_t y = 2 * cy + r;
        if (y < fdy && !held(right, y) && (right + 1 >= fdx || !held(right + 1, y))) wR += t.right[kOf(right, y)];
        const int32_t x = 2 * cx + r;
        if (x < fdx && !held(x, bottom) && (bottom + 1 >= fdy || !held(x, bottom + 1))) wD += t.down[kOf(x, bottom)];
    }
    const bool fixed = nfree == 0u && nheld > 0u;
    const uint32_t lc = l + 1u;
    const float scale = lc == 1u ? 1.0f / COARSEN : 1.0f;
    const uint32_t at = mgAt(a, lc, cx, cy);
    a.mg[at] = make_float4(0.0f, fixed ? 0.0f : rhs, wR * scale, wD * scale);
    a.mg[at + 1u] = make_float4(edge * scale, fixed ? 1.0f : 0.0f, 0.0f, 0.0f);
}

// down (restricts: the residual into level l + 1) or smoothTile (level 0
// the coarsest: its correction into its potential), on the square at
// (ox, oy), by NT threads
template <int T, int NT>
__device__ void downTile(const AirDev &a, AirTile<T> &t, uint32_t l, int32_t ox, int32_t oy, int restricts) {
    constexpr int SPAN = AirTile<T>::SPAN;
    const int32_t cx0 = ox - HALO, cy0 = oy - HALO;
    const int32_t dx = (int32_t)a.w[l], dy = (int32_t)a.h[l];
    loadTile<T, NT, false>(a, t, l, cx0, cy0);
    for (int i = 0; i < AIR_SMOOTH; i++) {
        sweepTile<T, NT>(t, 0, cx0, cy0, dx, dy);
        sweepTile<T, NT>(t, 1, cx0, cy0, dx, dy);
    }
    // storeTile (:1328-1339)
    for (int q = (int)threadIdx.x; q < T * T; q += NT) {
        const int32_t x = ox + q % T, y = oy + q / T;
        const int k = (q / T + HALO) * SPAN + q % T + HALO;
        if (x >= dx || y >= dy) continue;
        // level 0's into an array of its own, held cells' none (the reference
        // keeps it in cg's A p word, none there: cgStart's, cgApply's)
        if (restricts && l == 0u) a.pre0[(uint32_t)y * a.W + (uint32_t)x] = t.x[k];
        if (!(t.diag[k] >= 0.0f)) continue;
        if (!restricts) a.mg[mgAt(a, l, x, y)].x = t.x[k];
        else if (l != 0u) a.mg[mgAt(a, l, x, y) + 1u].z = t.x[k];
    }
    if (restricts) {
        const int32_t cw = (int32_t)a.w[l + 1u], ch = (int32_t)a.h[l + 1u];
        for (int q = (int)threadIdx.x; q < (T / 2) * (T / 2); q += NT) {
            const int32_t cx = ox / 2 + q % (T / 2), cy = oy / 2 + q / (T / 2);
            if (cx < cw && cy < ch) restrictCell(a, t, l, cx0, cy0, cx, cy);
        }
    }
}

// up, on the square at (ox, oy); with rho (level 0), cgRho's partials of
// the square's 8 x 8 groups, r . z with z as this pass leaves it, by the
// groups' tree, a wavefront a group
template <int T, int NT>
__device__ void upTile(const AirDev &a, AirTile<T> &t, uint32_t l, int32_t ox, int32_t oy, int rho) {
    constexpr int SPAN = AirTile<T>::SPAN;
    const int32_t cx0 = ox - HALO, cy0 = oy - HALO;
    const int32_t dx = (int32_t)a.w[l], dy = (int32_t)a.h[l];
    loadTile<T, NT, true>(a, t, l, cx0, cy0);
    for (int i = 0; i < AIR_SMOOTH; i++) {
        sweepTile<T, NT>(t, 1, cx0, cy0, dx, dy);
        sweepTile<T, NT>(t, 0, cx0, cy0, dx, dy);
    }
    for (int q = (int)threadIdx.x; q < T * T; q += NT) {
        const int32_t x = ox + q % T, y = oy + q / T;
        const int k = (q / T + HALO) * SPAN + q % T + HALO;
        if (x < dx && y < dy && t.diag[k] >= 0.0f) {
            if (l == 0u) a.z0[(uint32_t)y * a.W + (uint32_t)x] = t.x[k];
            else a.mg[mgAt(a, l, x, y)].x = t.x[k];
        }
    }
    if (!rho) return;
    const uint32_t lane = threadIdx.x & 63u, gh = (a.H + 7u) / 8u;
    for (uint32_t q = threadIdx.x / 64u; q < (uint32_t)(T / 8) * (T / 8); q += NT / 64) {
        const uint32_t gx = (uint32_t)ox / 8u + q % (T / 8), gy = (uint32_t)oy / 8u + q / (T / 8);
        if (gx >= a.gx || gy >= gh) continue;
        const uint32_t x = gx * 8u + lane % 8u, y = gy * 8u + lane / 8u;
        float v = 0.0f;
        if (x < a.W && y < a.H) {
            const int k = ((int32_t)y - cy0) * SPAN + ((int32_t)x - cx0);
            // a held cell keeps the potential it had (none: cgUpdate's, cgStart's)
            const float z = t.diag[k] >= 0.0f ? t.x[k] : a.z0[y * a.W + x];
            v = a.cg[y * a.W + x].y * z;
        }
        v = groupTree(v);
        if (lane == 0u) sums(a)[4u + gy * a.gx + gx] = v;
    }
}

// A level pass's square for this workgroup.  Workgroups go to the 8 XCDs in
// turn (the b-th to XCD b % 8), each XCD with an L2 of its own; so the
// squares are dealt out as 8 bands of consecutive squares, one an XCD, so
// that neighbours, whose halos overlap, share an L2.  (The last n % 8, if
// any, keep their own order: a one-to-one map of the squares.)
template <int T>
__device__ __forceinline__ int2 squareOrigin() {
    const uint32_t n = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x, full = n / 8u * 8u;
    const uint32_t q = b < full ? (b % 8u) * (n / 8u) + b / 8u : b;
    return make_int2((int32_t)(q % gridDim.x) * T, (int32_t)(q / gridDim.x) * T);
}

template <int T>
__global__ void __launch_bounds__(256) k_airDown(AirDev a, uint32_t l, int restricts, uint32_t cycle) {
    if (airConverged(a, cycle)) return;
    __shared__ AirTile<T> t;
    const int2 o = squareOrigin<T>();
    downTile<T, 256>(a, t, l, o.x, o.y, restricts);
}

template <int T>
__global__ void __launch_bounds__(256) k_airUp(AirDev a, uint32_t l, uint32_t cycle, int rho) {
    if (airConverged(a, cycle)) return;
    __shared__ AirTile<T> t;
    const int2 o = squareOrigin<T>();
    upTile<T, 256>(a, t, l, o.x, o.y, rho);
}

// solveCoarsest (:1482-1535): symmetric SOR in LDS, a cell a thread (the
// first 256 of the workgroup)
__device__ void coarsestBody(const AirDev &a, float *X) {
    const uint32_t l = a.L - 1u, dx = a.w[l], dy = a.h[l], n = min(dx * dy, (uint32_t)AIR_COARSEST);
    const uint32_t j = threadIdx.x;
    // only the potentials change from sweep to sweep: they go in LDS (X,
    // a float a cell), the cell's weights, rhs and diagonal stay in
    // registers, each the value the reference's sweep recomputes (the same
    // operations in the same order)
    const int32_t cx = (int32_t)(j % dx), cy = (int32_t)(j / dx);
    float4 m = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float lz = 0.0f, uw = 0.0f, diag = 1.0f;
    bool free_ = false;
    if (j < n) {
        m = a.mg[a.base[l] + 2u * j];
        const float4 b = a.mg[a.base[l] + 2u * j + 1u];
        X[j] = m.x;
        free_ = !(b.y > 0.0f);
        if (cx > 0) lz = a.mg[a.base[l] + 2u * (j - 1u)].z;
        if (cy > 0) uw = a.mg[a.base[l] + 2u * (j - dx)].w;
        diag = m.z + m.w + b.x + anchorAt(a, l, cx, cy);
        if (cx > 0) diag += lz;
        if (cy > 0) diag += uw;
    }
    __syncthreads();
    for (int sweep = 0; sweep < 2 * COARSE_SWEEPS; sweep++) {
        const int color = sweep < COARSE_SWEEPS ? (sweep & 1) : ((2 * COARSE_SWEEPS - 1 - sweep) & 1);
        if (j < n && ((cx + cy) & 1) == color && free_) {
            float sum = 0.0f;
            if ((uint32_t)cx + 1u < dx) sum += m.z * X[j + 1u];
            if ((uint32_t)cy + 1u < dy) sum += m.w * X[j + dx];
            if (cx > 0) sum += lz * X[j - 1u];
            if (cy > 0) sum += uw * X[j - dx];
            // mix(e1, e2, t) = e1 (1 - t) + e2 t
            const float xn = (sum - m.y) / diag;
            X[j] = X[j] * (1.0f - COARSE_OMEGA) + xn * COARSE_OMEGA;
        }
        __syncthreads();
    }
    if (j < n) {
        a.mg[a.base[l] + 2u * j].x = X[j];
        if (l == 0u) a.z0[j] = X[j];
    }
}
__global__ void __launch_bounds__(256) k_airCoarsest(AirDev a, uint32_t cycle) {
    if (airConverged(a, cycle)) return;
    __shared__ float X[AIR_COARSEST];
    coarsestBody(a, X);
}

// The levels from l0 (the first that fits a T x T square) to the coarsest
// and back, by one workgroup: each level's down as one square, the
// coarsest's SOR, each level's up; a level's results go to memory before
// the next phase reads them, as between the separate passes.  LDS
// (dynamic, COARSE_FUSED_BYTES): the tile, then the coarsest's cells.
#define COARSE_FUSED_T 64
#define COARSE_FUSED_NT 1024
#define COARSE_FUSED_BYTES (sizeof(AirTile<COARSE_FUSED_T>) + AIR_COARSEST * sizeof(float))
__global__ void __launch_bounds__(COARSE_FUSED_NT) k_airCoarse(AirDev a, uint32_t l0, uint32_t cycle) {
    if (airConverged(a, cycle)) return;
    extern __shared__ float4 smem[];
    AirTile<COARSE_FUSED_T> &t = *(AirTile<COARSE_FUSED_T> *)smem;
    float *X = (float *)((char *)smem + sizeof(AirTile<COARSE_FUSED_T>));
    const uint32_t deepest = a.L - 1u;
    for (uint32_t l = l0; l < deepest; l++) {
        downTile<COARSE_FUSED_T, COARSE_FUSED_NT>(a, t, l, 0, 0, 1);
        __syncthreads();
    }
    coarsestBody(a, X);
    __syncthreads();
    for (int32_t l = (int32_t)deepest - 1; l >= (int32_t)l0; l--) {
        upTile<COARSE_FUSED_T, COARSE_FUSED_NT>(a, t, (uint32_t)l, 0, 0, l == 0 ? 1 : 0);
        __syncthreads();
    }
}

// cgRho (:1099-1112) and cgApply (:1141-1157): a wavefront a group of 8 x 8
// level-0 cells, its partial by the group's tree into the sums
__global__ void __launch_bounds__(256) k_airRho(AirDev a, uint32_t cycle) {
    if (airConverged(a, cycle)) return;
    const uint32_t g = blockIdx.x * 4u + threadIdx.x / 64u, lane = threadIdx.x & 63u;
    if (g >= a.groups) return;
    const uint32_t x = (g % a.gx) * 8u + lane % 8u, y = (g / a.gx) * 8u + lane / 8u;
    float v = 0.0f;
    if (x < a.W && y < a.H) v = a.cg[y * a.W + x].y * a.mg[mgAt(a, 0u, (int32_t)x, (int32_t)y)].x;
    v = groupTree(v);
    if (lane == 0u) sums(a)[4u + g] = v;
}
// cgDirection (:1128-1139) and cgApply (:1141-1157) in one pass: each cell's
// p = z + beta p (none where held) from its old p, its neighbours' likewise
// (the same operands, so the same bits), A p as applyA has it; p into an
// array of its own (the neighbours still read the old one), which cgUpdate
// writes into cg's p
__device__ __forceinline__ float newP(float z, float4 w, float pold, float beta) {
    return w.w > 0.0f ? 0.0f : z + beta * pold;
}
__global__ void __launch_bounds__(256) k_airApply(AirDev a, uint32_t cycle) {
    if (airConverged(a, cycle)) return;
    // a wavefront's 8 x 8 group and the ring around it, 10 x 10 cells, in
    // LDS: the packed words and the new p of each (none off the grid)
    __shared__ float4 sw[4][100];
    __shared__ float sp[4][100];
    const uint32_t wave = threadIdx.x / 64u, g = blockIdx.x * 4u + wave, lane = threadIdx.x & 63u;
    const int32_t gx0 = (int32_t)((g % a.gx) * 8u) - 1, gy0 = (int32_t)((g / a.gx) * 8u) - 1;
    const float beta = sums(a)[2];
    for (uint32_t k = lane; k < 100u; k += 64u) {
        const int32_t cx = gx0 + (int32_t)(k % 10u), cy = gy0 + (int32_t)(k / 10u);
        const bool in = g < a.groups && cx >= 0 && cy >= 0 && cx < (int32_t)a.W && cy < (int32_t)a.H;
        const uint32_t ci = in ? (uint32_t)cy * a.W + (uint32_t)cx : 0u;
        const float4 w = a.w0[ci];
        const float z = a.z0[ci], po = a.cg[ci].z;
        sw[wave][k] = w;
        sp[wave][k] = in ? newP(z, w, po, beta) : 0.0f;
    }
    __syncthreads();
    const uint32_t x = (g % a.gx) * 8u + lane % 8u, y = (g / a.gx) * 8u + lane / 8u;
    float v = 0.0f;
    if (g < a.groups && x < a.W && y < a.H) {
        const int32_t xi = (int32_t)x, yi = (int32_t)y;
        const uint32_t i = y * a.W + x, k = (lane / 8u + 1u) * 10u + lane % 8u + 1u;
        const bool r_ = x + 1u < a.W, d_ = y + 1u < a.H, l_ = x > 0u, u_ = y > 0u;
        const float4 w = sw[wave][k], wl = sw[wave][k - 1u], wu = sw[wave][k - 10u];
        const float pc = sp[wave][k], pr = sp[wave][k + 1u], pd = sp[wave][k + 10u];
        const float pl = sp[wave][k - 1u], pu = sp[wave][k - 10u];
        float q = 0.0f;
        if (!(w.w > 0.0f)) {
            float diag = w.x + w.y + w.z + anchorAt(a, 0u, xi, yi);
            float sum = 0.0f;
            if (r_) sum += w.x * pr;
            if (d_) sum += w.y * pd;
            if (l_) {
                sum += wl.x * pl;
                diag += wl.x;
            }
            if (u_) {
                sum += wu.y * pu;
                diag += wu.y;
            }
            q = diag * pc - sum;
        }
        a.pnew[i] = pc;
        a.cg[i].w = q;
        v = pc * q;
    }
    v = groupTree(v);
    if (lane == 0u && g < a.groups) sums(a)[4u + g] = v;
}

// sumAll (:1081-1097): 256 lanes each its strided share, then halving
__device__ float sumAll(const AirDev &a) {
    __shared__ float wide[256];
    const uint32_t li = threadIdx.x;
    const float *s = sums(a) + 4;
    // each lane's share in order, its loads 16 at a time in flight
    float v = 0.0f;
    for (uint32_t k0 = li; k0 < a.groups; k0 += 16u * 256u) {
        float h[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint32_t k = k0 + 256u * (uint32_t)j;
            h[j] = k < a.groups ? s[k] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (k0 + 256u * (uint32_t)j < a.groups) v += h[j];
    }
    wide[li] = v;
    __syncthreads();
    if (li < 128u) wide[li] += wide[li + 128u];
    __syncthreads();
    if (li < 64u) wide[li] += wide[li + 64u];
    __syncthreads();
    float w = li < 64u ? wide[li] : 0.0f;
    w = groupTree(w);
    return w;   // lane 0's is the sum
}
// cgBeta (:1114-1126)
__global__ void __launch_bounds__(256) k_airBeta(AirDev a, uint32_t cycle) {
    if (airConverged(a, cycle)) return;
    const float rho = sumAll(a);
    if (threadIdx.x == 0u) {
        float *s = sums(a);
        const float before = s[0];
        s[2] = (cycle == 0u || before == 0.0f) ? 0.0f : rho / before;
        s[0] = rho;
    }
}
// cgAlpha (:1159-1170)
__global__ void __launch_bounds__(256) k_airAlpha(AirDev a, uint32_t cycle) {
    if (airConverged(a, cycle)) return;
    const float pq = sumAll(a);
    if (threadIdx.x == 0u) {
        float *s = sums(a);
        s[1] = pq <= 0.0f ? 0.0f : s[0] / pq;
    }
}

// cgUpdate (:1172-1194)
__global__ void __launch_bounds__(256) k_airUpdate(AirDev a, uint32_t cycle) {
    const bool conv = airConverged(a, cycle);
    if (blockIdx.x == 0u && threadIdx.x == 0u) {
        a.status[1u + (cycle + 1u) % 3u] = 0u;
        if (!conv) a.status[4] += 1u;
    }
    if (conv) return;
    uint32_t most = 0u;
    const float alpha = sums(a)[1];
    const uint32_t n = a.W * a.H, stride = gridDim.x * 256u;
    // UPDATE_HELD cells a thread at a time, their loads made together
    for (uint32_t i0 = blockIdx.x * 256u + threadIdx.x; i0 < n; i0 += UPDATE_HELD * stride) {
        float4 v[UPDATE_HELD];
#pragma unroll
        for (int j = 0; j < UPDATE_HELD; j++) {
            const uint32_t i = i0 + (uint32_t)j * stride;
            v[j] = a.cg[i < n ? i : 0u];
        }
#pragma unroll
        for (int j = 0; j < UPDATE_HELD; j++) {
            const uint32_t i = i0 + (uint32_t)j * stride;
            if (i >= n) continue;
            const float p = a.pnew[i];   // (cgDirection's, from cgApply)
            const float r = v[j].y - alpha * v[j].w;
            a.cg[i] = make_float4(v[j].x + alpha * p, r, p, v[j].w);
            a.rhs0[i] = -r;
            if (a.L == 1u) *(float2 *)&a.mg[mgAt(a, 0u, (int32_t)(i % a.W), (int32_t)(i / a.W))] = make_float2(0.0f, -r);
            most = max(most, __float_as_uint(fabsf(r)));
        }
    }
    groupMax(a.most, most);
}

// The level-0 records' potential and rhs as the last cgUpdate (or cgStart)
// left them: none, and -r
__global__ void __launch_bounds__(256) k_airFinish(AirDev a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.W * a.H || a.L == 1u) return;
    *(float2 *)&a.mg[mgAt(a, 0u, (int32_t)(i % a.W), (int32_t)(i / a.W))] = make_float2(0.0f, -a.cg[i].y);
}

// ------------------------------------------------------------ the C ABI

int wig_fail_message(int code, const char *msg);
extern "C" int wig_order_streams(hipStream_t from, hipStream_t to);   // (wig_api.hip)

struct WigAir {
    AirDev d;
    uint64_t mgBytes, cgBytes;
    hipStream_t stream = nullptr, own = nullptr;
    uint32_t tile[AIR_LEVELS] = {32, 16, 8, 8, 8, 8, 8, 8, 8};   // (measured at 1024 x 1024)
    int coarse = 0;        // the levels that fit COARSE_FUSED_T in one kernel
    int timing = 0;
    double ms = 0.0;
    uint32_t calls = 0;
    hipEvent_t t0 = nullptr, t1 = nullptr;
};

static int airFail(int code, const char *what, hipError_t e) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    return wig_fail_message(code, buf);
}
#define AIRCHECK(x)                                                 \
    do {                                                            \
        hipError_t e_ = (x);                                        \
        if (e_ != hipSuccess) return airFail(WIG_E_DEVICE, #x, e_); \
    } while (0)

// The grid's levels (js/sim.js:712-719): halved, rounding up, until a level
// holds AIR_COARSEST cells or fewer, AIR_LEVELS at most; their bases in
// airProjection, and the 8 x 8 groups.  The cells of every level, or 0 if
// the coarsest of AIR_LEVELS would hold more than AIR_COARSEST.
static uint64_t airLevels(AirDev &d, uint32_t width, uint32_t height) {
    d.W = width;
    d.H = height;
    d.L = 1;
    d.w[0] = width;
    d.h[0] = height;
    while (d.L < AIR_LEVELS && (uint64_t)d.w[d.L - 1] * d.h[d.L - 1] > AIR_COARSEST) {
        d.w[d.L] = (d.w[d.L - 1] + 1u) / 2u;
        d.h[d.L] = (d.h[d.L - 1] + 1u) / 2u;
        d.L++;
    }
    if ((uint64_t)d.w[d.L - 1] * d.h[d.L - 1] > AIR_COARSEST) return 0;
    uint64_t cells = 0;
    for (uint32_t l = 0; l < d.L; l++) {
        d.base[l] = (uint32_t)(2u * cells);
        cells += (uint64_t)d.w[l] * d.h[l];
    }
    for (uint32_t l = d.L; l < AIR_LEVELS; l++) d.w[l] = d.h[l] = d.base[l] = 0;
    d.gx = (width + 7u) / 8u;
    d.groups = d.gx * ((height + 7u) / 8u);
    return cells;
}
// The working arrays after `most`: w0 (16 B a cell), then pre0, pnew, rhs0,
// z0 (4 B each)
static void airWork(AirDev &d, void *work) {
    const uint64_t n = (uint64_t)d.W * d.H;
    d.w0 = (float4 *)work;
    d.pre0 = (float *)(d.w0 + n);
    d.pnew = d.pre0 + n;
    d.rhs0 = d.pnew + n;
    d.z0 = d.rhs0 + n;
}

extern "C" WigAir *wig_air_create(uint32_t width, uint32_t height) {
    if (width == 0 || height == 0 || width > 16384 || height > 16384) {
        wig_fail_message(WIG_E_INVALID, "wig_air_create: sides must be 1..16384");
        return nullptr;
    }
    WigAir *a = new WigAir();
    AirDev &d = a->d;
    const uint64_t cells = airLevels(d, width, height);
    if (cells == 0) {
        wig_fail_message(WIG_E_INVALID, "wig_air_create: the coarsest of 9 levels would hold more than 256 cells");
        delete a;
        return nullptr;
    }
    a->mgBytes = 32u * cells;
    a->cgBytes = 16u * ((uint64_t)width * height + (4u + d.groups + 3u) / 4u);
    void *mg = nullptr, *cg = nullptr, *st = nullptr, *mean = nullptr;
    hipError_t e = hipMalloc(&mg, a->mgBytes);
    if (e == hipSuccess) e = hipMalloc(&cg, a->cgBytes);
    if (e == hipSuccess) e = hipMalloc(&st, 32);
    if (e == hipSuccess) e = hipMalloc(&mean, 4ull * width * height);
    void *most = nullptr, *work = nullptr;
    if (e == hipSuccess) e = hipMalloc(&most, 4ull * CELL_GROUPS);
    if (e == hipSuccess) e = hipMalloc(&work, 32ull * width * height);
    if (e == hipSuccess) e = hipStreamCreate(&a->own);
    if (e == hipSuccess) e = hipMemsetAsync(work, 0, 32ull * width * height, a->own);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_airCoarse, hipFuncAttributeMaxDynamicSharedMemorySize, (int)COARSE_FUSED_BYTES);
    if (e == hipSuccess) e = hipMemsetAsync(mg, 0, a->mgBytes, a->own);
    if (e == hipSuccess) e = hipMemsetAsync(cg, 0, a->cgBytes, a->own);
    if (e == hipSuccess) e = hipMemsetAsync(st, 0, 32, a->own);
    if (e == hipSuccess) e = hipMemsetAsync(mean, 0, 4ull * width * height, a->own);
    if (e == hipSuccess) e = hipStreamSynchronize(a->own);
    if (e != hipSuccess) {
        airFail(e == hipErrorOutOfMemory ? WIG_E_NOMEM : WIG_E_DEVICE, "wig_air_create", e);
        if (mg) (void)hipFree(mg);
        if (cg) (void)hipFree(cg);
        if (st) (void)hipFree(st);
        if (mean) (void)hipFree(mean);
        if (most) (void)hipFree(most);
        if (work) (void)hipFree(work);
        if (a->own) (void)hipStreamDestroy(a->own);
        delete a;
        return nullptr;
    }
    d.mg = (float4 *)mg;
    d.cg = (float4 *)cg;
    d.status = (uint32_t *)st;
    d.mean = (const float *)mean;
    d.most = (uint32_t *)most;
    airWork(d, work);
    a->stream = a->own;
    return a;
}

extern "C" int wig_air_destroy(WigAir *a) {
    if (!a) return wig_fail_message(WIG_E_INVALID, "wig_air_destroy: NULL");
    (void)hipStreamSynchronize(a->stream);
    (void)hipFree(a->d.mg);
    (void)hipFree(a->d.cg);
    (void)hipFree(a->d.status);
    (void)hipFree((void *)a->d.mean);
    (void)hipFree(a->d.most);
    (void)hipFree(a->d.w0);
    if (a->t0) (void)hipEventDestroy(a->t0);
    if (a->t1) (void)hipEventDestroy(a->t1);
    if (a->own) (void)hipStreamDestroy(a->own);
    delete a;
    return WIG_OK;
}

extern "C" int wig_air_layout(const WigAir *a, WigAirLayout *out) {
    if (!a || !out) return wig_fail_message(WIG_E_INVALID, "wig_air_layout: NULL");
    memset(out, 0, sizeof *out);
    out->width = a->d.W;
    out->height = a->d.H;
    out->levels = a->d.L;
    out->groups = a->d.groups;
    for (uint32_t l = 0; l < a->d.L; l++) {
        out->level_width[l] = a->d.w[l];
        out->level_height[l] = a->d.h[l];
        out->level_base[l] = a->d.base[l];
    }
    out->part_bytes[WIG_AIR_PART_MG] = a->mgBytes;
    out->part_bytes[WIG_AIR_PART_CG] = a->cgBytes;
    out->part_bytes[WIG_AIR_PART_STATUS] = 32;
    out->part_bytes[WIG_AIR_PART_MEAN] = 4ull * a->d.W * a->d.H;
    return WIG_OK;
}

static void *airPart(WigAir *a, int part, uint64_t *bytes) {
    switch (part) {
    case WIG_AIR_PART_MG: *bytes = a->mgBytes; return a->d.mg;
    case WIG_AIR_PART_CG: *bytes = a->cgBytes; return a->d.cg;
    case WIG_AIR_PART_STATUS: *bytes = 32; return a->d.status;
    case WIG_AIR_PART_MEAN: *bytes = 4ull * a->d.W * a->d.H; return (void *)a->d.mean;
    default: return nullptr;
    }
}

extern "C" int wig_air_upload(WigAir *a, int part, const void *src, size_t bytes) {
    if (!a || !src) return wig_fail_message(WIG_E_INVALID, "wig_air_upload: NULL");
    uint64_t n = 0;
    void *p = airPart(a, part, &n);
    if (!p || bytes != n) return wig_fail_message(WIG_E_INVALID, "wig_air_upload: no such part, or not its size");
    AIRCHECK(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, a->stream));
    AIRCHECK(hipStreamSynchronize(a->stream));
    return WIG_OK;
}

extern "C" int wig_air_download(WigAir *a, int part, void *dst, size_t bytes) {
    if (!a || !dst) return wig_fail_message(WIG_E_INVALID, "wig_air_download: NULL");
    uint64_t n = 0;
    void *p = airPart(a, part, &n);
    if (!p || bytes != n) return wig_fail_message(WIG_E_INVALID, "wig_air_download: no such part, or not its size");
    AIRCHECK(hipMemcpyAsync(dst, p, bytes, hipMemcpyDeviceToHost, a->stream));
    AIRCHECK(hipStreamSynchronize(a->stream));
    return WIG_OK;
}

extern "C" int wig_air_device_ptr(WigAir *a, int part, void **ptr) {
    if (!a || !ptr) return wig_fail_message(WIG_E_INVALID, "wig_air_device_ptr: NULL");
    uint64_t n = 0;
    *ptr = airPart(a, part, &n);
    return *ptr ? WIG_OK : wig_fail_message(WIG_E_INVALID, "wig_air_device_ptr: no such part");
}

extern "C" int wig_air_set_stream(WigAir *a, void *stream) {
    if (!a) return wig_fail_message(WIG_E_INVALID, "wig_air_set_stream: NULL");
    hipStream_t next = stream ? (hipStream_t)stream : a->own;
    int st = wig_order_streams(a->stream, next);
    if (st < 0) return st;
    a->stream = next;
    return WIG_OK;
}

extern "C" int wig_air_set_coarse(WigAir *a, int on) {
    if (!a) return wig_fail_message(WIG_E_INVALID, "wig_air_set_coarse: NULL");
    a->coarse = on ? 1 : 0;
    return WIG_OK;
}

extern "C" int wig_air_set_tile(WigAir *a, uint32_t level, uint32_t side) {
    if (!a || (side != 8 && side != 16 && side != 32) || level >= a->d.L)
        return wig_fail_message(WIG_E_INVALID, "wig_air_set_tile: a level of the grid, side 8, 16 or 32");
    a->tile[level] = side;
    return WIG_OK;
}

template <int T>
static void launchLevel(const WigAir *a, bool up, uint32_t l, int restricts, uint32_t cycle) {
    const AirDev &d = a->d;
    const dim3 grid((d.w[l] + T - 1) / T, (d.h[l] + T - 1) / T);
    if (up) k_airUp<T><<<grid, 256, 0, a->stream>>>(d, l, cycle, l == 0u ? 1 : 0);
    else k_airDown<T><<<grid, 256, 0, a->stream>>>(d, l, restricts, cycle);
}

extern "C" int wig_air_project(WigAir *a, uint32_t cycles) {
    if (!a) return wig_fail_message(WIG_E_INVALID, "wig_air_project: NULL");
    if (cycles > 4096) return wig_fail_message(WIG_E_INVALID, "wig_air_project: at most 4096 iterations");
    const AirDev &d = a->d;
    const uint32_t n = d.W * d.H, deepest = d.L - 1u;
    const uint32_t cellBlocks = (n + 255u) / 256u, groupBlocks = (d.groups + 3u) / 4u;
    const uint32_t cellGrid = cellBlocks < CELL_GROUPS ? cellBlocks : CELL_GROUPS;
    if (a->timing) {
        if (!a->t0) AIRCHECK(hipEventCreate(&a->t0));
        if (!a->t1) AIRCHECK(hipEventCreate(&a->t1));
        AIRCHECK(hipEventRecord(a->t0, a->stream));
    }
    auto level = [&](bool up, uint32_t l, int restricts, uint32_t k) {
        if (a->tile[l] == 32) launchLevel<32>(a, up, l, restricts, k);
        else if (a->tile[l] == 8) launchLevel<8>(a, up, l, restricts, k);
        else launchLevel<16>(a, up, l, restricts, k);
    };
    // the first level that fits one square, if the fused kernel takes it
    uint32_t l0 = deepest;
    if (a->coarse && deepest > 0)
        for (uint32_t l = 0; l < deepest; l++)
            if (d.w[l] <= COARSE_FUSED_T && d.h[l] <= COARSE_FUSED_T) {
                l0 = l;
                break;
            }
    k_airClear<<<1, 64, 0, a->stream>>>(d);
    k_airStart<<<cellGrid, 256, 0, a->stream>>>(d);
    k_airSlot<<<1, 1024, 0, a->stream>>>(d, 3u, cellGrid, 0u, 1);
    for (uint32_t k = 0; k < cycles; k++) {
        if (l0 < deepest) {
            for (uint32_t l = 0; l < l0; l++) level(false, l, 1, k);
            k_airCoarse<<<1, COARSE_FUSED_NT, COARSE_FUSED_BYTES, a->stream>>>(d, l0, k);
            for (int32_t l = (int32_t)l0 - 1; l >= 0; l--) level(true, (uint32_t)l, 0, k);
        } else {
            for (uint32_t l = 0; l < deepest; l++) level(false, l, 1, k);
            if (deepest == 0) level(false, 0, 0, k);
            k_airCoarsest<<<1, 256, 0, a->stream>>>(d, k);
            for (int32_t l = (int32_t)deepest - 1; l >= 0; l--) level(true, (uint32_t)l, 0, k);
        }
        if (deepest == 0) k_airRho<<<groupBlocks, 256, 0, a->stream>>>(d, k);   // (else in level 0's up)
        k_airBeta<<<1, 256, 0, a->stream>>>(d, k);
        k_airApply<<<groupBlocks, 256, 0, a->stream>>>(d, k);
        k_airAlpha<<<1, 256, 0, a->stream>>>(d, k);
        k_airUpdate<<<cellGrid, 256, 0, a->stream>>>(d, k);
        k_airSlot<<<1, 1024, 0, a->stream>>>(d, 1u + k % 3u, cellGrid, k, 0);
    }
    k_airFinish<<<cellBlocks, 256, 0, a->stream>>>(d);
    AIRCHECK(hipGetLastError());
    if (a->timing) {
        AIRCHECK(hipEventRecord(a->t1, a->stream));
        AIRCHECK(hipEventSynchronize(a->t1));
        float ms = 0.0f;
        AIRCHECK(hipEventElapsedTime(&ms, a->t0, a->t1));
        a->ms += ms;
        a->calls++;
    }
    return WIG_OK;
}

// ------------------------------------------------------------ allocation-free projection
//
// wig_air_project_static: the projection of wig_air_project on buffers the
// caller holds, for callers that cannot allocate (code under HIP graph
// capture, serving loops that forbid hipMalloc).  Its working arrays (the
// workgroups' maxima and level 0's packed words, 4 KB + 32 B a cell) live in
// `scratch`, one array per device for the life of the process.  Host code
// never passes `scratch` itself to a kernel: in host code the name stands
// for a host-side shadow, not for the device's copy.  Each device's address
// comes from hipGetSymbolAddress, taken once per device with that device
// current, and only that address goes into kernel arguments.
#define AIR_STATIC_MOST_BYTES (4ull * CELL_GROUPS)
#define AIR_STATIC_BYTES (AIR_STATIC_MOST_BYTES + 32ull * WIG_AIR_STATIC_MAX_CELLS)
static __device__ __align__(256) unsigned char scratch[AIR_STATIC_BYTES];

#define AIR_STATIC_DEVICES 64
static void *g_scratchAt[AIR_STATIC_DEVICES];
static std::once_flag g_scratchOnce[AIR_STATIC_DEVICES];
static hipError_t g_scratchErr[AIR_STATIC_DEVICES];

// zeroes n 32-bit words (a fill kernel, not hipMemsetAsync, so that a
// caller's graph capture records a kernel node)
__global__ void __launch_bounds__(256) k_airZero(uint32_t *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) p[i] = 0u;
}

extern "C" int wig_air_project_static(void *mg, void *cg, void *status, const void *mean, uint32_t width,
                                      uint32_t height, uint32_t cycles, void *stream) {
    if (!mg || !cg || !status || !mean) return wig_fail_message(WIG_E_INVALID, "wig_air_project_static: NULL buffer");
    if (width == 0 || height == 0 || (uint64_t)width * height > WIG_AIR_STATIC_MAX_CELLS)
        return wig_fail_message(WIG_E_INVALID,
                                "wig_air_project_static: the grid must hold 1 .. WIG_AIR_STATIC_MAX_CELLS (262144) cells");
    if (cycles > 4096) return wig_fail_message(WIG_E_INVALID, "wig_air_project_static: at most 4096 iterations");
    int dev = 0;
    AIRCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= AIR_STATIC_DEVICES)
        return wig_fail_message(WIG_E_INVALID, "wig_air_project_static: device ordinal past 63");
    // the device's copy of the array, once per device (this device is current)
    std::call_once(g_scratchOnce[dev], [dev] {
        g_scratchErr[dev] = hipGetSymbolAddress(&g_scratchAt[dev], HIP_SYMBOL(scratch));
    });
    if (g_scratchErr[dev] != hipSuccess) return airFail(WIG_E_DEVICE, "wig_air_project_static: hipGetSymbolAddress", g_scratchErr[dev]);
    unsigned char *at = (unsigned char *)g_scratchAt[dev];
    WigAir a;   // a view on the caller's buffers and the array: nothing allocated
    AirDev &d = a.d;
    if (airLevels(d, width, height) == 0)
        return wig_fail_message(WIG_E_INVALID, "wig_air_project_static: the coarsest of 9 levels would hold more than 256 cells");
    d.mg = (float4 *)mg;
    d.cg = (float4 *)cg;
    d.status = (uint32_t *)status;
    d.mean = (const float *)mean;
    d.most = (uint32_t *)at;
    airWork(d, at + AIR_STATIC_MOST_BYTES);
    a.stream = (hipStream_t)stream;
    // what the last call left: cleared (held cells' pre0 must read none)
    const uint64_t words = (AIR_STATIC_MOST_BYTES + 32ull * width * height) / 4u;
    const uint64_t blocks = std::min<uint64_t>((words + 255u) / 256u, 4096u);
    k_airZero<<<(uint32_t)blocks, 256, 0, a.stream>>>((uint32_t *)at, words);
    AIRCHECK(hipGetLastError());
    return wig_air_project(&a, cycles);
}

extern "C" int wig_air_sync(WigAir *a) {
    if (!a) return wig_fail_message(WIG_E_INVALID, "wig_air_sync: NULL");
    AIRCHECK(hipStreamSynchronize(a->stream));
    return WIG_OK;
}

extern "C" int wig_air_time(WigAir *a, int on) {
    if (!a) return wig_fail_message(WIG_E_INVALID, "wig_air_time: NULL");
    a->timing = on;
    a->ms = 0.0;
    a->calls = 0;
    return WIG_OK;
}

extern "C" int wig_air_elapsed(WigAir *a, double *ms, uint32_t *calls) {
    if (!a || !ms || !calls) return wig_fail_message(WIG_E_INVALID, "wig_air_elapsed: NULL");
    *ms = a->ms;
    *calls = a->calls;
    return WIG_OK;
}
