/*
 * wig.node — a thin Node-API addon over libwig.so's C ABI (include/wig.h),
 * so the JavaScript host (liquid.js, the liquid part of js/sim.js's
 * Simulation) calls the HIP liquid step as the reference's host encodes its
 * WGSL passes (js/sim.js:2466-2612).  Buffers cross as ArrayBuffers and
 * typed arrays, copied; errors throw with wig_last_error()'s message.
 */
#define NAPI_VERSION 8
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/wig.h"

#define CALL(env, x)                                                     \
    do {                                                                 \
        if ((x) != napi_ok) {                                            \
            napi_throw_error((env), NULL, "Node-API call failed: " #x);  \
            return NULL;                                                 \
        }                                                                \
    } while (0)

static napi_value throw_wig(napi_env env, int status) {
    char msg[600];
    snprintf(msg, sizeof msg, "wig error %d: %s", status, wig_last_error());
    napi_throw_error(env, NULL, msg);
    return NULL;
}

static void finalize_world(napi_env env, void *data, void *hint) {
    (void)env;
    (void)hint;
    /* a world the script did not destroy: freed with its handle */
    WigWorld **slot = (WigWorld **)data;
    if (*slot) wig_destroy(*slot);
    *slot = NULL;
    free(slot);
}

static WigWorld *world_of(napi_env env, napi_value v) {
    void *data = NULL;
    if (napi_get_value_external(env, v, &data) != napi_ok || !data || !*(WigWorld **)data) {
        napi_throw_type_error(env, NULL, "not a live world handle");
        return NULL;
    }
    return *(WigWorld **)data;
}

static int bytes_of(napi_env env, napi_value v, void **data, size_t *len) {
    bool is;
    if (napi_is_arraybuffer(env, v, &is) == napi_ok && is) return napi_get_arraybuffer_info(env, v, data, len) == napi_ok;
    if (napi_is_typedarray(env, v, &is) == napi_ok && is) {
        napi_typedarray_type t;
        size_t n, off;
        napi_value ab;
        if (napi_get_typedarray_info(env, v, &t, &n, data, &ab, &off) != napi_ok) return 0;
        size_t el = (t == napi_int8_array || t == napi_uint8_array || t == napi_uint8_clamped_array) ? 1
                  : (t == napi_int16_array || t == napi_uint16_array) ? 2
                  : (t == napi_float64_array || t == napi_bigint64_array || t == napi_biguint64_array) ? 8 : 4;
        *len = n * el;
        return 1;
    }
    return 0;
}

static uint32_t arg_u32(napi_env env, napi_value v) {
    uint32_t x = 0;
    napi_get_value_uint32(env, v, &x);
    return x;
}

/* create(width, height, coarseLevels) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 2) { napi_throw_type_error(env, NULL, "create(width, height, coarseLevels?)"); return NULL; }
    WigWorld *w = wig_create(arg_u32(env, argv[0]), arg_u32(env, argv[1]), argc > 2 ? arg_u32(env, argv[2]) : 0u, 0u);
    if (!w) return throw_wig(env, WIG_E_INVALID);
    WigWorld **slot = (WigWorld **)malloc(sizeof *slot);
    *slot = w;
    napi_value h;
    CALL(env, napi_create_external(env, slot, finalize_world, NULL, &h));
    return h;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    void *data = NULL;
    CALL(env, napi_get_value_external(env, argv[0], &data));
    WigWorld **slot = (WigWorld **)data;
    if (*slot) wig_destroy(*slot);
    *slot = NULL;
    return NULL;
}

/* layout(h) -> {width, height, cells, faces, blocks, capacity, levels, partBytes: [...]} */
static napi_value js_layout(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1], o, v, arr;
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigWorld *w = world_of(env, argv[0]);
    if (!w) return NULL;
    WigLayoutInfo L;
    int st = wig_layout(w, &L);
    if (st < 0) return throw_wig(env, st);
    CALL(env, napi_create_object(env, &o));
#define FIELD(name, val) CALL(env, napi_create_uint32(env, (val), &v)); CALL(env, napi_set_named_property(env, o, name, v))
    FIELD("width", L.width);
    FIELD("height", L.height);
    FIELD("cells", L.cells);
    FIELD("faces", L.faces);
    FIELD("blocks", L.blocks);
    FIELD("capacity", L.capacity);
    FIELD("levels", L.levels);
#undef FIELD
    CALL(env, napi_create_array_with_length(env, WIG_PART_COUNT, &arr));
    for (uint32_t p = 0; p < WIG_PART_COUNT; p++) {
        CALL(env, napi_create_double(env, (double)L.part_bytes[p], &v));
        CALL(env, napi_set_element(env, arr, p, v));
    }
    CALL(env, napi_set_named_property(env, o, "partBytes", arr));
    return o;
}

/* setKinds(h, Uint8Array(256)) */
static napi_value js_set_kinds(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigWorld *w = world_of(env, argv[0]);
    void *data;
    size_t len;
    if (!w) return NULL;
    if (!bytes_of(env, argv[1], &data, &len) || len != 256) { napi_throw_type_error(env, NULL, "kinds: 256 bytes"); return NULL; }
    int st = wig_set_kinds(w, (const uint8_t *)data);
    return st < 0 ? throw_wig(env, st) : NULL;
}

/* setElementProps(h, Float32Array(256 * 6)): viscosity, viscosityE,
 * viscosityT, tension, wetting, breaks a row (WigElementProps) */
static napi_value js_set_element_props(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigWorld *w = world_of(env, argv[0]);
    void *data;
    size_t len;
    if (!w) return NULL;
    if (!bytes_of(env, argv[1], &data, &len) || len != 256 * sizeof(WigElementProps)) {
        napi_throw_type_error(env, NULL, "element props: 256 rows of 6 floats");
        return NULL;
    }
    int st = wig_set_element_props(w, (const WigElementProps *)data);
    return st < 0 ? throw_wig(env, st) : NULL;
}

/* upload(h, part, bytes) / download(h, part, bytes) */
static napi_value js_copy(napi_env env, napi_callback_info info, int up) {
    size_t argc = 3;
    napi_value argv[3];
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigWorld *w = world_of(env, argv[0]);
    void *data;
    size_t len;
    if (!w) return NULL;
    if (!bytes_of(env, argv[2], &data, &len)) { napi_throw_type_error(env, NULL, "an ArrayBuffer or typed array"); return NULL; }
    int part = (int)arg_u32(env, argv[1]);
    int st = up ? wig_upload(w, part, data, len) : wig_download(w, part, data, len);
    return st < 0 ? throw_wig(env, st) : NULL;
}
static napi_value js_upload(napi_env env, napi_callback_info info) { return js_copy(env, info, 1); }
static napi_value js_download(napi_env env, napi_callback_info info) { return js_copy(env, info, 0); }

/* fluidStep(h, params: ArrayBuffer(80), ctl: ArrayBuffer(88)) -> status */
static napi_value js_fluid_step(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3], r;
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigWorld *w = world_of(env, argv[0]);
    void *p, *c;
    size_t pl, cl;
    if (!w) return NULL;
    if (!bytes_of(env, argv[1], &p, &pl) || pl != sizeof(WigParams) || !bytes_of(env, argv[2], &c, &cl) || cl != sizeof(WigStepCtl)) {
        napi_throw_type_error(env, NULL, "fluidStep(h, params (80 bytes), ctl (88 bytes))");
        return NULL;
    }
    WigParams P;
    WigStepCtl C;
    memcpy(&P, p, sizeof P);
    memcpy(&C, c, sizeof C);
    int st = wig_fluid_step(w, &P, &C);
    if (st < 0) return throw_wig(env, st);
    CALL(env, napi_create_int32(env, st, &r));
    return r;
}

/* report(h) -> Uint32Array(9): FLUID_STATS, MOST_CYCLES cleared after */
static napi_value js_report(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1], ab, ta;
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigWorld *w = world_of(env, argv[0]);
    if (!w) return NULL;
    void *data;
    CALL(env, napi_create_arraybuffer(env, 4 * WIG_STATS, &data, &ab));
    int st = wig_report(w, (uint32_t *)data);
    if (st < 0) return throw_wig(env, st);
    CALL(env, napi_create_typedarray(env, napi_uint32_array, WIG_STATS, ab, 0, &ta));
    return ta;
}

/* reportAsync(h) -> ticket: the report asked for without waiting (wig_report_async) */
static napi_value js_report_async(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1], r;
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigWorld *w = world_of(env, argv[0]);
    if (!w) return NULL;
    uint64_t ticket = 0;
    int st = wig_report_async(w, &ticket);
    if (st < 0) return throw_wig(env, st);
    CALL(env, napi_create_double(env, (double)ticket, &r));
    return r;
}

/* reportTake(h, ticket, wait) -> Uint32Array(9), or null if it has not landed
 * and wait is false (wig_report_take) */
static napi_value js_report_take(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3], ab, ta;
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigWorld *w = world_of(env, argv[0]);
    if (!w) return NULL;
    double ticket = 0.0;
    bool wait = true;
    CALL(env, napi_get_value_double(env, argv[1], &ticket));
    if (argc > 2) CALL(env, napi_get_value_bool(env, argv[2], &wait));
    void *data;
    CALL(env, napi_create_arraybuffer(env, 4 * WIG_STATS, &data, &ab));
    int st = wig_report_take(w, (uint64_t)ticket, (uint32_t *)data, wait ? 1 : 0);
    if (st == WIG_NOT_READY) {
        napi_value nul;
        CALL(env, napi_get_null(env, &nul));
        return nul;
    }
    if (st < 0) return throw_wig(env, st);
    CALL(env, napi_create_typedarray(env, napi_uint32_array, WIG_STATS, ab, 0, &ta));
    return ta;
}

/* solveStats(h) -> Float64Array(6) [rz, alpha, beta, pq, rr, bb] plus .cycles */
static napi_value js_solve_stats(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1], o, v;
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigWorld *w = world_of(env, argv[0]);
    if (!w) return NULL;
    WigSolveStats s;
    int st = wig_solve_stats(w, &s);
    if (st < 0) return throw_wig(env, st);
    CALL(env, napi_create_object(env, &o));
#define D(name, val) CALL(env, napi_create_double(env, (val), &v)); CALL(env, napi_set_named_property(env, o, name, v))
    D("rz", s.rz);
    D("alpha", s.alpha);
    D("beta", s.beta);
    D("pq", s.pq);
    D("rr", s.rr);
    D("bb", s.bb);
    D("cycles", (double)s.cycles);
    D("converged", (double)s.converged);
#undef D
    return o;
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigWorld *w = world_of(env, argv[0]);
    if (!w) return NULL;
    int st = wig_sync(w);
    return st < 0 ? throw_wig(env, st) : NULL;
}

static napi_value js_device_ok(napi_env env, napi_callback_info info) {
    (void)info;
    napi_value r;
    CALL(env, napi_get_boolean(env, wig_device_ok() != 0, &r));
    return r;
}

/* ---- the air's projection (wig_air_*) ---- */
static void finalize_air(napi_env env, void *data, void *hint) {
    (void)env;
    (void)hint;
    WigAir **slot = (WigAir **)data;
    if (*slot) wig_air_destroy(*slot);
    *slot = NULL;
    free(slot);
}

static WigAir *air_of(napi_env env, napi_value v) {
    void *data = NULL;
    if (napi_get_value_external(env, v, &data) != napi_ok || !data || !*(WigAir **)data) {
        napi_throw_type_error(env, NULL, "not a live air handle");
        return NULL;
    }
    return *(WigAir **)data;
}

/* airCreate(width, height) -> handle */
static napi_value js_air_create(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 2) { napi_throw_type_error(env, NULL, "airCreate(width, height)"); return NULL; }
    WigAir *a = wig_air_create(arg_u32(env, argv[0]), arg_u32(env, argv[1]));
    if (!a) return throw_wig(env, WIG_E_INVALID);
    WigAir **slot = (WigAir **)malloc(sizeof *slot);
    *slot = a;
    napi_value h;
    CALL(env, napi_create_external(env, slot, finalize_air, NULL, &h));
    return h;
}

static napi_value js_air_destroy(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    void *data = NULL;
    CALL(env, napi_get_value_external(env, argv[0], &data));
    WigAir **slot = (WigAir **)data;
    if (*slot) wig_air_destroy(*slot);
    *slot = NULL;
    return NULL;
}

/* airLayout(h) -> {width, height, levels, groups, levelWidth, levelHeight, partBytes} */
static napi_value js_air_layout(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1], o, v, arr;
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigAir *a = air_of(env, argv[0]);
    if (!a) return NULL;
    WigAirLayout L;
    int st = wig_air_layout(a, &L);
    if (st < 0) return throw_wig(env, st);
    CALL(env, napi_create_object(env, &o));
#define FIELD(name, val) CALL(env, napi_create_uint32(env, (val), &v)); CALL(env, napi_set_named_property(env, o, name, v))
    FIELD("width", L.width);
    FIELD("height", L.height);
    FIELD("levels", L.levels);
    FIELD("groups", L.groups);
#undef FIELD
#define ARRAY(name, n, get) \
    CALL(env, napi_create_array_with_length(env, (n), &arr)); \
    for (uint32_t k = 0; k < (n); k++) { CALL(env, napi_create_double(env, (double)(get), &v)); CALL(env, napi_set_element(env, arr, k, v)); } \
    CALL(env, napi_set_named_property(env, o, name, arr))
    ARRAY("levelWidth", L.levels, L.level_width[k]);
    ARRAY("levelHeight", L.levels, L.level_height[k]);
    ARRAY("partBytes", WIG_AIR_PART_COUNT, L.part_bytes[k]);
#undef ARRAY
    return o;
}

/* airUpload(h, part, bytes) / airDownload(h, part, bytes) */
static napi_value js_air_copy(napi_env env, napi_callback_info info, int up) {
    size_t argc = 3;
    napi_value argv[3];
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigAir *a = air_of(env, argv[0]);
    void *data;
    size_t len;
    if (!a) return NULL;
    if (!bytes_of(env, argv[2], &data, &len)) { napi_throw_type_error(env, NULL, "an ArrayBuffer or typed array"); return NULL; }
    int part = (int)arg_u32(env, argv[1]);
    int st = up ? wig_air_upload(a, part, data, len) : wig_air_download(a, part, data, len);
    return st < 0 ? throw_wig(env, st) : NULL;
}
static napi_value js_air_upload(napi_env env, napi_callback_info info) { return js_air_copy(env, info, 1); }
static napi_value js_air_download(napi_env env, napi_callback_info info) { return js_air_copy(env, info, 0); }

/* airProject(h, cycles): enqueued */
static napi_value js_air_project(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigAir *a = air_of(env, argv[0]);
    if (!a) return NULL;
    int st = wig_air_project(a, argc > 1 ? arg_u32(env, argv[1]) : 64u);
    return st < 0 ? throw_wig(env, st) : NULL;
}

static napi_value js_air_sync(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    WigAir *a = air_of(env, argv[0]);
    if (!a) return NULL;
    int st = wig_air_sync(a);
    return st < 0 ? throw_wig(env, st) : NULL;
}

static napi_value init(napi_env env, napi_value exports) {
    napi_property_descriptor props[] = {
        {"create", NULL, js_create, NULL, NULL, NULL, napi_default, NULL},
        {"destroy", NULL, js_destroy, NULL, NULL, NULL, napi_default, NULL},
        {"layout", NULL, js_layout, NULL, NULL, NULL, napi_default, NULL},
        {"setKinds", NULL, js_set_kinds, NULL, NULL, NULL, napi_default, NULL},
        {"setElementProps", NULL, js_set_element_props, NULL, NULL, NULL, napi_default, NULL},
        {"upload", NULL, js_upload, NULL, NULL, NULL, napi_default, NULL},
        {"download", NULL, js_download, NULL, NULL, NULL, napi_default, NULL},
        {"fluidStep", NULL, js_fluid_step, NULL, NULL, NULL, napi_default, NULL},
        {"report", NULL, js_report, NULL, NULL, NULL, napi_default, NULL},
        {"reportAsync", NULL, js_report_async, NULL, NULL, NULL, napi_default, NULL},
        {"reportTake", NULL, js_report_take, NULL, NULL, NULL, napi_default, NULL},
        {"solveStats", NULL, js_solve_stats, NULL, NULL, NULL, napi_default, NULL},
        {"sync", NULL, js_sync, NULL, NULL, NULL, napi_default, NULL},
        {"deviceOk", NULL, js_device_ok, NULL, NULL, NULL, napi_default, NULL},
        {"airCreate", NULL, js_air_create, NULL, NULL, NULL, napi_default, NULL},
        {"airDestroy", NULL, js_air_destroy, NULL, NULL, NULL, napi_default, NULL},
        {"airLayout", NULL, js_air_layout, NULL, NULL, NULL, napi_default, NULL},
        {"airUpload", NULL, js_air_upload, NULL, NULL, NULL, napi_default, NULL},
        {"airDownload", NULL, js_air_download, NULL, NULL, NULL, napi_default, NULL},
        {"airProject", NULL, js_air_project, NULL, NULL, NULL, napi_default, NULL},
        {"airSync", NULL, js_air_sync, NULL, NULL, NULL, napi_default, NULL},
    };
    if (napi_define_properties(env, exports, sizeof props / sizeof props[0], props) != napi_ok) return NULL;
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
