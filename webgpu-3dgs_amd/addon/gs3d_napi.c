/*
 * gs3d_napi.c — Node N-API binding of libgs3d.so (include/gs3d.h).
 *
 * A thin C layer: every function converts JS arguments (numbers, typed
 * arrays, strings), calls one C-ABI entry point and turns a non-zero return
 * into a thrown JS Error carrying gs3d_last_error(). The context is a JS
 * external whose finalizer destroys it. Built against N-API 8 headers; the
 * napi_* symbols resolve from the node binary at load time.
 */
#define _GNU_SOURCE /* MAP_POPULATE */
#include <fcntl.h>
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "gs3d.h"

#define NAPI_CALL(env, call)                                                    \
    do {                                                                        \
        if ((call) != napi_ok) {                                                \
            napi_throw_error((env), NULL, "gs3d addon: N-API call failed: " #call); \
            return NULL;                                                        \
        }                                                                       \
    } while (0)

#define ARGS(n)                                                                 \
    size_t argc = (n);                                                          \
    napi_value argv[(n) > 0 ? (n) : 1];                                         \
    NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL))

static napi_value throw_rc(napi_env env, gs3d_ctx* c, int rc, const char* what) {
    char msg[768];
    snprintf(msg, sizeof msg, "%s failed (%d): %s", what, rc, c ? gs3d_last_error(c) : "");
    char code[16];
    snprintf(code, sizeof code, "%d", rc);
    napi_throw_error(env, code, msg);
    return NULL;
}

#define CHECK(env, c, expr, what)                       \
    do {                                                \
        int rc_ = (expr);                               \
        if (rc_ != GS3D_OK) return throw_rc(env, c, rc_, what); \
    } while (0)

static void finalize_ctx(napi_env env, void* data, void* hint) {
    (void)env;
    (void)hint;
    gs3d_ctx** slot = (gs3d_ctx**)data;
    if (*slot) gs3d_destroy(*slot);
    free(slot);
}

static gs3d_ctx* get_ctx(napi_env env, napi_value v) {
    void* data = NULL;
    if (napi_get_value_external(env, v, &data) != napi_ok || !data || !*(gs3d_ctx**)data) {
        napi_throw_type_error(env, NULL, "gs3d: invalid or destroyed context handle");
        return NULL;
    }
    return *(gs3d_ctx**)data;
}

static int get_u32(napi_env env, napi_value v, uint32_t* out) {
    return napi_get_value_uint32(env, v, out) == napi_ok;
}

/* Argument conversions that throw a TypeError naming the call on failure
 * (a wrong type never silently becomes 0). */
static int arg_u32(napi_env env, napi_value v, uint32_t* out, const char* usage) {
    double d;
    if (napi_get_value_double(env, v, &d) != napi_ok || !(d >= 0.0) || d > 4294967295.0 || d != (double)(uint32_t)d) {
        napi_throw_type_error(env, NULL, usage);
        return 0;
    }
    *out = (uint32_t)d;
    return 1;
}

static int arg_i32(napi_env env, napi_value v, int32_t* out, const char* usage) {
    if (napi_get_value_int32(env, v, out) != napi_ok) {
        napi_throw_type_error(env, NULL, usage);
        return 0;
    }
    return 1;
}

static int is_nullish(napi_env env, napi_value v) {
    napi_valuetype t;
    return napi_typeof(env, v, &t) == napi_ok && (t == napi_undefined || t == napi_null);
}

/* A byte count or offset: a non-negative integer below 2^53. */
static int arg_size(napi_env env, napi_value v, size_t* out, const char* usage) {
    double d;
    if (napi_get_value_double(env, v, &d) != napi_ok || !(d >= 0.0) || d > 9007199254740991.0 ||
        d != (double)(uint64_t)d) {
        napi_throw_type_error(env, NULL, usage);
        return 0;
    }
    *out = (size_t)d;
    return 1;
}

static int arg_bool(napi_env env, napi_value v, bool* out, const char* usage) {
    if (napi_get_value_bool(env, v, out) != napi_ok) {
        napi_throw_type_error(env, NULL, usage);
        return 0;
    }
    return 1;
}

/* Data pointer + byte length of a TypedArray / ArrayBuffer / DataView;
 * null and undefined give NULL. */
static int get_bytes(napi_env env, napi_value v, void** data, size_t* bytes) {
    napi_valuetype t;
    if (napi_typeof(env, v, &t) != napi_ok) return 0;
    if (t == napi_null || t == napi_undefined) {
        *data = NULL;
        *bytes = 0;
        return 1;
    }
    bool is;
    if (napi_is_typedarray(env, v, &is) == napi_ok && is) {
        napi_typedarray_type tt;
        size_t len, off;
        napi_value ab;
        if (napi_get_typedarray_info(env, v, &tt, &len, data, &ab, &off) != napi_ok) return 0;
        size_t el = 1;
        switch (tt) {
            case napi_int16_array: case napi_uint16_array: el = 2; break;
            case napi_int32_array: case napi_uint32_array: case napi_float32_array: el = 4; break;
            case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: el = 8; break;
            default: el = 1;
        }
        *bytes = len * el;
        return 1;
    }
    if (napi_is_arraybuffer(env, v, &is) == napi_ok && is)
        return napi_get_arraybuffer_info(env, v, data, bytes) == napi_ok;
    return 0;
}

static napi_value make_u32(napi_env env, uint32_t v) {
    napi_value r;
    napi_create_uint32(env, v, &r);
    return r;
}

static napi_value make_typed(napi_env env, napi_typedarray_type t, const void* src, size_t count,
                             size_t el) {
    void* data;
    napi_value ab, out;
    if (napi_create_arraybuffer(env, count * el, &data, &ab) != napi_ok) return NULL;
    if (count) memcpy(data, src, count * el);
    if (napi_create_typedarray(env, t, count, ab, 0, &out) != napi_ok) return NULL;
    return out;
}

static napi_value pair(napi_env env, uint32_t a, uint32_t b) {
    napi_value arr;
    napi_create_array_with_length(env, 2, &arr);
    napi_set_element(env, arr, 0, make_u32(env, a));
    napi_set_element(env, arr, 1, make_u32(env, b));
    return arr;
}

/* create(device) -> external */
static napi_value js_create(napi_env env, napi_callback_info info) {
    ARGS(1);
    int32_t dev = 0;
    if (argc > 0 && !is_nullish(env, argv[0]) && !arg_i32(env, argv[0], &dev, "create(device?)")) return NULL;
    gs3d_ctx** slot = (gs3d_ctx**)malloc(sizeof *slot);
    if (!slot) {
        napi_throw_error(env, NULL, "create: out of memory");
        return NULL;
    }
    gs3d_ctx* c = NULL;
    const int rc = gs3d_create(dev, &c);
    if (rc != GS3D_OK) {
        free(slot);
        return throw_rc(env, NULL, rc, "gs3d_create");
    }
    *slot = c;
    napi_value ext;
    NAPI_CALL(env, napi_create_external(env, slot, finalize_ctx, NULL, &ext));
    return ext;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
    ARGS(1);
    void* data = NULL;
    NAPI_CALL(env, napi_get_value_external(env, argv[0], &data));
    gs3d_ctx** slot = (gs3d_ctx**)data;
    if (slot && *slot) {
        gs3d_destroy(*slot);
        *slot = NULL;
    }
    return NULL;
}

static napi_value js_last_error(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    napi_value s;
    NAPI_CALL(env, napi_create_string_utf8(env, gs3d_last_error(c), NAPI_AUTO_LENGTH, &s));
    return s;
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    CHECK(env, c, gs3d_sync(c), "gs3d_sync");
    return NULL;
}

static napi_value js_device_limits(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint32_t lim[6];
    CHECK(env, c, gs3d_device_limits(c, lim), "gs3d_device_limits");
    return make_typed(env, napi_uint32_array, lim, 6, 4);
}

/* setSplats(ctx, pos, scale, rot, color, fDc|null, shRest|null, shDegree, n) */
static napi_value js_set_splats(napi_env env, napi_callback_info info) {
    ARGS(9);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    void* p[6];
    size_t b[6];
    for (int k = 0; k < 6; k++)
        if (!get_bytes(env, argv[1 + k], &p[k], &b[k])) {
            napi_throw_type_error(env, NULL, "setSplats: expected typed arrays");
            return NULL;
        }
    int32_t deg = 0;
    uint32_t n = 0;
    if (!arg_i32(env, argv[7], &deg, "setSplats: shDegree must be an integer") ||
        !arg_u32(env, argv[8], &n, "setSplats: n must be a non-negative integer"))
        return NULL;
    for (int k = 0; k < 4; k++)
        if (b[k] < (size_t)n * 16) {
            napi_throw_range_error(env, NULL, "setSplats: vec4 array shorter than 16*n bytes");
            return NULL;
        }
    if (deg > 0) {
        const size_t rest = (size_t)3 * ((deg + 1) * (deg + 1) - 1);
        if (b[4] < (size_t)n * 12 || b[5] < (size_t)n * rest * 4) {
            napi_throw_range_error(env, NULL, "setSplats: f_dc / sh_rest shorter than n rows");
            return NULL;
        }
    }
    CHECK(env, c, gs3d_set_splats(c, p[0], p[1], p[2], p[3], p[4], p[5], deg, n), "gs3d_set_splats");
    return NULL;
}

static napi_value js_set_tile_size(napi_env env, napi_callback_info info) {
    ARGS(3);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint32_t x = 0, y = 0, out[2];
    if (!arg_u32(env, argv[1], &x, "setTileSize(ctx, x, y)") || !arg_u32(env, argv[2], &y, "setTileSize(ctx, x, y)"))
        return NULL;
    CHECK(env, c, gs3d_set_tile_size(c, x, y, out), "gs3d_set_tile_size");
    return pair(env, out[0], out[1]);
}

static napi_value js_get_tile_size(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint32_t out[2];
    CHECK(env, c, gs3d_get_tile_size(c, out), "gs3d_get_tile_size");
    return pair(env, out[0], out[1]);
}

static napi_value js_get_tile_count(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint32_t out[2];
    CHECK(env, c, gs3d_get_tile_count(c, out), "gs3d_get_tile_count");
    return pair(env, out[0], out[1]);
}

/* rebin(ctx, camera Float32Array(72), width, height, flags) */
static napi_value js_rebin(napi_env env, napi_callback_info info) {
    ARGS(5);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    void* cam;
    size_t cb;
    if (!get_bytes(env, argv[1], &cam, &cb) || cb < sizeof(gs3d_camera)) {
        napi_throw_type_error(env, NULL, "rebin: camera must be a Float32Array of 72 floats");
        return NULL;
    }
    uint32_t w = 0, h = 0, flags = 0;
    static const char* usage = "rebin(ctx, camera, width, height, flags?)";
    if (!arg_u32(env, argv[2], &w, usage) || !arg_u32(env, argv[3], &h, usage)) return NULL;
    if (argc > 4 && !is_nullish(env, argv[4]) && !arg_u32(env, argv[4], &flags, usage)) return NULL;
    CHECK(env, c, gs3d_rebin(c, (const gs3d_camera*)cam, w, h, flags), "gs3d_rebin");
    return NULL;
}

static napi_value js_ref_count(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint64_t k = 0;
    CHECK(env, c, gs3d_ref_count(c, &k), "gs3d_ref_count");
    napi_value r;
    napi_create_double(env, (double)k, &r);
    return r;
}

/* read(ctx, bufferId) -> ArrayBuffer */
static napi_value js_read(napi_env env, napi_callback_info info) {
    ARGS(2);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    int32_t id = 0;
    if (!arg_i32(env, argv[1], &id, "read(ctx, bufferId)")) return NULL;
    size_t size = 0;
    CHECK(env, c, gs3d_read(c, id, NULL, 0, &size), "gs3d_read");
    void* data;
    napi_value ab;
    NAPI_CALL(env, napi_create_arraybuffer(env, size, &data, &ab));
    CHECK(env, c, gs3d_read(c, id, data, size, &size), "gs3d_read");
    return ab;
}

/* readTileCounts(ctx) -> Uint32Array | null (null on any failure) */
static napi_value js_read_tile_counts(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    napi_value nul;
    napi_get_null(env, &nul);
    if (!c) return NULL;
    size_t n = 0;
    if (gs3d_read_tile_counts(c, NULL, 0, &n) != GS3D_OK) return nul;
    void* data;
    napi_value ab, out;
    NAPI_CALL(env, napi_create_arraybuffer(env, n * 4, &data, &ab));
    if (gs3d_read_tile_counts(c, (uint32_t*)data, n, &n) != GS3D_OK) return nul;
    NAPI_CALL(env, napi_create_typedarray(env, napi_uint32_array, n, ab, 0, &out));
    return out;
}

static napi_value js_stage_times(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    float t[GS3D_STAGE__COUNT];
    CHECK(env, c, gs3d_stage_times_us(c, t, GS3D_STAGE__COUNT), "gs3d_stage_times_us");
    return make_typed(env, napi_float32_array, t, GS3D_STAGE__COUNT, 4);
}

static napi_value js_stage_times_avg(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    float t[GS3D_STAGE__COUNT];
    uint64_t frames = 0;
    CHECK(env, c, gs3d_stage_times_avg_us(c, t, GS3D_STAGE__COUNT, &frames), "gs3d_stage_times_avg_us");
    return make_typed(env, napi_float32_array, t, GS3D_STAGE__COUNT, 4);
}

static napi_value js_profiler_reset(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    CHECK(env, c, gs3d_profiler_reset(c), "gs3d_profiler_reset");
    return NULL;
}

static napi_value js_stage_label(napi_env env, napi_callback_info info) {
    ARGS(1);
    int32_t i = 0;
    if (!arg_i32(env, argv[0], &i, "stageLabel(stage)")) return NULL;
    napi_value s;
    NAPI_CALL(env, napi_create_string_utf8(env, gs3d_stage_label(i), NAPI_AUTO_LENGTH, &s));
    return s;
}

static int get_name(napi_env env, napi_value v, char* buf, size_t cap) {
    size_t len = 0;
    return napi_get_value_string_utf8(env, v, buf, cap, &len) == napi_ok && len > 0 && len < cap - 1;
}

static int arg_name(napi_env env, napi_value v, char* buf, size_t cap) {
    if (get_name(env, v, buf, cap)) return 1;
    napi_throw_type_error(env, NULL, "buffer name must be a non-empty string of at most 62 bytes");
    return 0;
}

static napi_value js_buffer_create(napi_env env, napi_callback_info info) {
    ARGS(3);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    char name[64];
    size_t bytes = 0;
    if (!arg_name(env, argv[1], name, sizeof name) ||
        !arg_size(env, argv[2], &bytes, "bufferCreate(ctx, name, bytes)"))
        return NULL;
    CHECK(env, c, gs3d_buffer_create(c, name, bytes), "gs3d_buffer_create");
    return NULL;
}

static napi_value js_buffer_resize(napi_env env, napi_callback_info info) {
    ARGS(3);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    char name[64];
    size_t bytes = 0;
    if (!arg_name(env, argv[1], name, sizeof name) ||
        !arg_size(env, argv[2], &bytes, "bufferResize(ctx, name, bytes)"))
        return NULL;
    CHECK(env, c, gs3d_buffer_resize(c, name, bytes), "gs3d_buffer_resize");
    return NULL;
}

/* bufferSize(ctx, name) -> number | undefined for an unknown name */
static napi_value js_buffer_size(napi_env env, napi_callback_info info) {
    ARGS(2);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    char name[64];
    if (!arg_name(env, argv[1], name, sizeof name)) return NULL;
    size_t bytes = 0;
    if (gs3d_buffer_size(c, name, &bytes) != GS3D_OK) return NULL;
    napi_value r;
    napi_create_double(env, (double)bytes, &r);
    return r;
}

/* bufferWrite(ctx, name, offset, data) -> bytes written (0 = skipped) */
static napi_value js_buffer_write(napi_env env, napi_callback_info info) {
    ARGS(4);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    char name[64];
    size_t off = 0;
    if (!arg_name(env, argv[1], name, sizeof name) ||
        !arg_size(env, argv[2], &off, "bufferWrite(ctx, name, offset, data)"))
        return NULL;
    void* data;
    size_t bytes;
    if (!get_bytes(env, argv[3], &data, &bytes)) {
        napi_throw_type_error(env, NULL, "bufferWrite: data must be a typed array or ArrayBuffer");
        return NULL;
    }
    size_t written = 0;
    int rc = gs3d_buffer_write(c, name, off, data, bytes, &written);
    if (rc != GS3D_OK && rc != GS3D_E_TOO_SMALL) return throw_rc(env, c, rc, "gs3d_buffer_write");
    napi_value r;
    napi_create_double(env, (double)written, &r);
    return r;
}

static napi_value js_buffer_read(napi_env env, napi_callback_info info) {
    ARGS(4);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    char name[64];
    size_t off = 0, bytes = 0;
    if (!arg_name(env, argv[1], name, sizeof name) ||
        !arg_size(env, argv[2], &off, "bufferRead(ctx, name, offset, bytes)") ||
        !arg_size(env, argv[3], &bytes, "bufferRead(ctx, name, offset, bytes)"))
        return NULL;
    void* data;
    napi_value ab;
    NAPI_CALL(env, napi_create_arraybuffer(env, bytes, &data, &ab));
    CHECK(env, c, gs3d_buffer_read(c, name, off, data, bytes), "gs3d_buffer_read");
    return ab;
}

static napi_value js_buffer_destroy(napi_env env, napi_callback_info info) {
    ARGS(2);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    char name[64];
    if (!arg_name(env, argv[1], name, sizeof name)) return NULL;
    CHECK(env, c, gs3d_buffer_destroy(c, name), "gs3d_buffer_destroy");
    return NULL;
}

/* bufferList(ctx) -> [{name, size}], largest first */
static napi_value js_buffer_list(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    size_t count = 0;
    CHECK(env, c, gs3d_buffer_list(c, NULL, NULL, 0, &count), "gs3d_buffer_list");
    char* names = (char*)calloc(count ? count : 1, 64);
    uint64_t* sizes = (uint64_t*)calloc(count ? count : 1, 8);
    int rc = gs3d_buffer_list(c, names, sizes, count, &count);
    napi_value arr;
    napi_create_array_with_length(env, count, &arr);
    for (size_t i = 0; rc == GS3D_OK && i < count; i++) {
        napi_value o, nm, sz;
        napi_create_object(env, &o);
        napi_create_string_utf8(env, names + 64 * i, NAPI_AUTO_LENGTH, &nm);
        napi_create_double(env, (double)sizes[i], &sz);
        napi_set_named_property(env, o, "name", nm);
        napi_set_named_property(env, o, "size", sz);
        napi_set_element(env, arr, (uint32_t)i, o);
    }
    free(names);
    free(sizes);
    if (rc != GS3D_OK) return throw_rc(env, c, rc, "gs3d_buffer_list");
    return arr;
}

/* setKernelVariant(ctx, kernel, param, value): PipelineManager.update with new
 * codeConstants (PipelineManager.ts:66-105). */
static napi_value js_set_kernel_variant(napi_env env, napi_callback_info info) {
    ARGS(4);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    char kernel[64], param[64];
    int32_t value = 0;
    if (!get_name(env, argv[1], kernel, sizeof kernel) || !get_name(env, argv[2], param, sizeof param) ||
        napi_get_value_int32(env, argv[3], &value) != napi_ok) {
        napi_throw_type_error(env, NULL, "setKernelVariant(ctx, kernel, param, value)");
        return NULL;
    }
    CHECK(env, c, gs3d_set_kernel_variant(c, kernel, param, value), "gs3d_set_kernel_variant");
    return NULL;
}

static napi_value js_get_kernel_variant(napi_env env, napi_callback_info info) {
    ARGS(3);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    char kernel[64], param[64];
    if (!get_name(env, argv[1], kernel, sizeof kernel) || !get_name(env, argv[2], param, sizeof param)) {
        napi_throw_type_error(env, NULL, "getKernelVariant(ctx, kernel, param)");
        return NULL;
    }
    int value = 0;
    CHECK(env, c, gs3d_get_kernel_variant(c, kernel, param, &value), "gs3d_get_kernel_variant");
    napi_value r;
    napi_create_int32(env, value, &r);
    return r;
}

/* launchGeometry(ctx | null, kernel, [x, y, z]) -> {workgroupSize: [3], dispatchSize: [3]}:
 * the launch shape libgs3d.so uses for one pass of `kernel` (gs3d_launch_geometry). */
static napi_value js_launch_geometry(napi_env env, napi_callback_info info) {
    ARGS(3);
    gs3d_ctx* c = NULL;
    if (!is_nullish(env, argv[0])) {
        c = get_ctx(env, argv[0]);
        if (!c) return NULL;
    }
    char kernel[64];
    uint64_t problem[3] = {1, 1, 1};
    bool is_array = false;
    uint32_t len = 0;
    if (!get_name(env, argv[1], kernel, sizeof kernel) || napi_is_array(env, argv[2], &is_array) != napi_ok ||
        !is_array || napi_get_array_length(env, argv[2], &len) != napi_ok || len < 1 || len > 3) {
        napi_throw_type_error(env, NULL, "launchGeometry(ctx | null, kernel, [x, y, z])");
        return NULL;
    }
    for (uint32_t i = 0; i < len; i++) {
        napi_value e;
        size_t v = 0;
        if (napi_get_element(env, argv[2], i, &e) != napi_ok ||
            !arg_size(env, e, &v, "launchGeometry(ctx | null, kernel, [x, y, z])"))
            return NULL;
        problem[i] = v;
    }
    uint32_t wg[3];
    uint64_t grid[3];
    const int rc = gs3d_launch_geometry(c, kernel, problem, wg, grid);
    if (rc != GS3D_OK) {
        char msg[160];
        snprintf(msg, sizeof msg, "gs3d_launch_geometry failed (%d): no pass named '%s'", rc, kernel);
        char code[16];
        snprintf(code, sizeof code, "%d", rc);
        napi_throw_error(env, code, msg);
        return NULL;
    }
    napi_value out, w, d;
    napi_create_object(env, &out);
    napi_create_array_with_length(env, 3, &w);
    napi_create_array_with_length(env, 3, &d);
    for (uint32_t i = 0; i < 3; i++) {
        napi_value x;
        napi_create_uint32(env, wg[i], &x);
        napi_set_element(env, w, i, x);
        napi_create_double(env, (double)grid[i], &x);
        napi_set_element(env, d, i, x);
    }
    napi_set_named_property(env, out, "workgroupSize", w);
    napi_set_named_property(env, out, "dispatchSize", d);
    return out;
}

/* referencePassStages(label) -> bit mask of the stages doing that reference pass's work. */
static napi_value js_reference_pass_stages(napi_env env, napi_callback_info info) {
    ARGS(1);
    char label[64];
    if (!get_name(env, argv[0], label, sizeof label)) {
        napi_throw_type_error(env, NULL, "referencePassStages(label)");
        return NULL;
    }
    return make_u32(env, gs3d_reference_pass_stages(label));
}

/* setProfileMask(ctx, mask): which stages a profiled rebin brackets with events. */
static napi_value js_set_profile_mask(napi_env env, napi_callback_info info) {
    ARGS(2);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint32_t mask = 0;
    if (!get_u32(env, argv[1], &mask)) {
        napi_throw_type_error(env, NULL, "setProfileMask(ctx, mask)");
        return NULL;
    }
    CHECK(env, c, gs3d_set_profile_mask(c, mask), "gs3d_set_profile_mask");
    return NULL;
}

/* sortPasses(ctx): bit offsets of the key-sort passes the last rebin ran. */
static napi_value js_sort_passes(napi_env env, napi_callback_info info) {
    ARGS(1);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint32_t n = 0, offs[16];
    CHECK(env, c, gs3d_sort_passes(c, &n, offs), "gs3d_sort_passes");
    return make_typed(env, napi_uint32_array, offs, n, 4);
}

/* setDebugMode(ctx, mode, debugRef) */
static napi_value js_set_debug_mode(napi_env env, napi_callback_info info) {
    ARGS(3);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint32_t mode = 0, ref = 256;
    if (!get_u32(env, argv[1], &mode) || (argc > 2 && !get_u32(env, argv[2], &ref))) {
        napi_throw_type_error(env, NULL, "setDebugMode(ctx, mode, debugRef)");
        return NULL;
    }
    CHECK(env, c, gs3d_set_debug_mode(c, mode, ref), "gs3d_set_debug_mode");
    return NULL;
}

/* composite(ctx, format, clear Float32Array(4) | null, framebuffer Uint8Array | null):
 * blends the last frame onto the framebuffer in place (W*H*4 bytes). */
static napi_value js_composite(napi_env env, napi_callback_info info) {
    ARGS(4);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    int32_t format = 0;
    void *clear, *fb;
    size_t cb, fbb;
    if (napi_get_value_int32(env, argv[1], &format) != napi_ok || !get_bytes(env, argv[2], &clear, &cb) ||
        !get_bytes(env, argv[3], &fb, &fbb) || (clear && cb < 16)) {
        napi_throw_type_error(env, NULL, "composite(ctx, format, clear Float32Array(4)|null, framebuffer Uint8Array|null)");
        return NULL;
    }
    CHECK(env, c, gs3d_composite(c, format, (const float*)clear, (uint8_t*)fb, fbb), "gs3d_composite");
    return NULL;
}

static napi_value f32_array(napi_env env, size_t count, float** data) {
    napi_value ab, out;
    if (napi_create_arraybuffer(env, count * 4, (void**)data, &ab) != napi_ok) return NULL;
    if (napi_create_typedarray(env, napi_float32_array, count, ab, 0, &out) != napi_ok) return NULL;
    return out;
}

static napi_value xyz_object(napi_env env, const double v[3]) {
    napi_value o, x;
    napi_create_object(env, &o);
    const char* names[3] = {"x", "y", "z"};
    for (int i = 0; i < 3; i++) {
        napi_create_double(env, v[i], &x);
        napi_set_named_property(env, o, names[i], x);
    }
    return o;
}

static napi_value parse_ply_bytes(napi_env env, const void* data, size_t bytes, int32_t threads);

/* parsePly(bytes | path, threads?) -> {splats: GaussianSplatData (+ fDc), bounds}:
 * the loader's result (SceneWorker.ts:336-405) parsed natively; throws the
 * loader's error message on a malformed file. A path string is memory-mapped. */
static napi_value js_parse_ply(napi_env env, napi_callback_info info) {
    ARGS(2);
    int32_t threads = 0;
    if (argc > 1 && !is_nullish(env, argv[1]) &&
        !arg_i32(env, argv[1], &threads, "parsePly(bytes: Buffer | ArrayBuffer | TypedArray | path, threads?)"))
        return NULL;
    napi_valuetype t;
    napi_typeof(env, argv[0], &t);
    if (t == napi_string) {
        char path[4096];
        size_t len = 0;
        if (napi_get_value_string_utf8(env, argv[0], path, sizeof path, &len) != napi_ok || len >= sizeof path - 1) {
            napi_throw_type_error(env, NULL, "parsePly: path too long");
            return NULL;
        }
        const int fd = open(path, O_RDONLY);
        struct stat st;
        if (fd < 0 || fstat(fd, &st) != 0) {
            char msg[4200];
            snprintf(msg, sizeof msg, "parsePly: cannot open %s", path);
            if (fd >= 0) close(fd);
            napi_throw_error(env, NULL, msg);
            return NULL;
        }
        if (st.st_size == 0) {
            close(fd);
            static const char empty = 0;
            return parse_ply_bytes(env, &empty, 0, threads);
        }
        void* map = mmap(NULL, (size_t)st.st_size, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
        close(fd);
        if (map == MAP_FAILED) {
            napi_throw_error(env, NULL, "parsePly: mmap failed");
            return NULL;
        }
        napi_value r = parse_ply_bytes(env, map, (size_t)st.st_size, threads);
        munmap(map, (size_t)st.st_size);
        return r;
    }
    void* data;
    size_t bytes;
    if (!get_bytes(env, argv[0], &data, &bytes) || !data) {
        napi_throw_type_error(env, NULL, "parsePly(bytes: Buffer | ArrayBuffer | TypedArray | path, threads?)");
        return NULL;
    }
    return parse_ply_bytes(env, data, bytes, threads);
}

static napi_value parse_ply_bytes(napi_env env, const void* data, size_t bytes, int32_t threads) {
    gs3d_ply_info pi;
    char err[512] = {0};
    if (gs3d_ply_probe(data, bytes, &pi, err, sizeof err) != GS3D_OK) {
        napi_throw_error(env, NULL, err);
        return NULL;
    }
    const size_t n = (size_t)pi.splat_count, F = pi.f_rest_count;
    float *pos, *sc, *rot, *col, *dc, *rest = NULL;
    napi_value jpos = f32_array(env, n * 4, &pos), jcol = f32_array(env, n * 4, &col);
    napi_value jsc = f32_array(env, n * 4, &sc), jrot = f32_array(env, n * 4, &rot);
    napi_value jdc = f32_array(env, n * 3, &dc), jrest;
    if (F) jrest = f32_array(env, n * F, &rest);
    else napi_get_null(env, &jrest);
    if (!jpos || !jcol || !jsc || !jrot || !jdc || !jrest) {
        napi_throw_error(env, NULL, "parsePly: out of memory");
        return NULL;
    }
    if (gs3d_ply_parse(data, bytes, pos, sc, rot, col, dc, rest, threads, &pi, err, sizeof err) != GS3D_OK) {
        napi_throw_error(env, NULL, err);
        return NULL;
    }
    napi_value splats, bounds, result, v;
    napi_create_object(env, &splats);
    napi_set_named_property(env, splats, "positions", jpos);
    napi_set_named_property(env, splats, "colors", jcol);
    napi_set_named_property(env, splats, "scales", jsc);
    napi_set_named_property(env, splats, "rotations", jrot);
    napi_set_named_property(env, splats, "fDc", jdc);
    napi_set_named_property(env, splats, "sphericalHarmonics", jrest);
    napi_create_int32(env, pi.sh_degree, &v);
    napi_set_named_property(env, splats, "sphericalHarmonicsDegree", v);
    napi_create_double(env, (double)n, &v);
    napi_set_named_property(env, splats, "splatCount", v);
    napi_create_object(env, &bounds);
    napi_set_named_property(env, bounds, "min", xyz_object(env, pi.bounds_min));
    napi_set_named_property(env, bounds, "max", xyz_object(env, pi.bounds_max));
    napi_create_object(env, &result);
    napi_set_named_property(env, result, "splats", splats);
    napi_set_named_property(env, result, "bounds", bounds);
    return result;
}

/* loadPly(ctx, path, threads?) -> {splatCount, sphericalHarmonicsDegree, fRestCount, bounds}:
 * parse natively and upload (SceneSyncer.setSplatData). */
static napi_value js_load_ply(napi_env env, napi_callback_info info) {
    ARGS(3);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    char path[4096];
    size_t len = 0;
    if (napi_get_value_string_utf8(env, argv[1], path, sizeof path, &len) != napi_ok || len >= sizeof path - 1) {
        napi_throw_type_error(env, NULL, "loadPly(ctx, path, threads?)");
        return NULL;
    }
    int32_t threads = 0;
    if (argc > 2 && !is_nullish(env, argv[2]) && !arg_i32(env, argv[2], &threads, "loadPly(ctx, path, threads?)"))
        return NULL;
    gs3d_ply_info pi;
    CHECK(env, c, gs3d_load_ply(c, path, threads, &pi), "gs3d_load_ply");
    napi_value r, v, bounds;
    napi_create_object(env, &r);
    napi_create_double(env, (double)pi.splat_count, &v);
    napi_set_named_property(env, r, "splatCount", v);
    napi_create_int32(env, pi.sh_degree, &v);
    napi_set_named_property(env, r, "sphericalHarmonicsDegree", v);
    napi_create_uint32(env, pi.f_rest_count, &v);
    napi_set_named_property(env, r, "fRestCount", v);
    napi_create_object(env, &bounds);
    napi_set_named_property(env, bounds, "min", xyz_object(env, pi.bounds_min));
    napi_set_named_property(env, bounds, "max", xyz_object(env, pi.bounds_max));
    napi_set_named_property(env, r, "bounds", bounds);
    return r;
}

/* views: Array<{camera: Float32Array(72), width, height}> -> malloc'd gs3d_view[] */
static gs3d_view* get_views(napi_env env, napi_value arr, uint32_t* count) {
    bool is_array = false;
    if (napi_is_array(env, arr, &is_array) != napi_ok || !is_array) return NULL;
    uint32_t n = 0;
    napi_get_array_length(env, arr, &n);
    if (n == 0) return NULL;
    gs3d_view* v = (gs3d_view*)calloc(n, sizeof(gs3d_view));
    if (!v) return NULL;
    for (uint32_t i = 0; i < n; i++) {
        napi_value e, cam, w, h;
        void* data;
        size_t bytes;
        if (napi_get_element(env, arr, i, &e) != napi_ok || napi_get_named_property(env, e, "camera", &cam) != napi_ok ||
            !get_bytes(env, cam, &data, &bytes) || !data || bytes < sizeof(gs3d_camera) ||
            napi_get_named_property(env, e, "width", &w) != napi_ok || !get_u32(env, w, &v[i].width) ||
            napi_get_named_property(env, e, "height", &h) != napi_ok || !get_u32(env, h, &v[i].height)) {
            free(v);
            return NULL;
        }
        memcpy(&v[i].camera, data, sizeof(gs3d_camera));
    }
    *count = n;
    return v;
}

/* rebinBatch(ctx, views, flags) */
static napi_value js_rebin_batch(napi_env env, napi_callback_info info) {
    ARGS(3);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint32_t n = 0, flags = 0;
    gs3d_view* v = get_views(env, argv[1], &n);
    if (!v) {
        napi_throw_type_error(env, NULL, "rebinBatch(ctx, [{camera: Float32Array(72), width, height}, ...], flags)");
        return NULL;
    }
    if (argc > 2) get_u32(env, argv[2], &flags);
    const int rc = gs3d_rebin_batch(c, v, n, flags);
    free(v);
    CHECK(env, c, rc, "gs3d_rebin_batch");
    return NULL;
}

/* batchRefCounts(ctx) -> Float64Array (K per view) */
static napi_value js_batch_ref_counts(napi_env env, napi_callback_info info) {
    ARGS(2);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint64_t k[GS3D_MAX_BATCH_VIEWS];
    uint32_t n = 0;
    if (!get_u32(env, argv[1], &n) || n == 0 || n > GS3D_MAX_BATCH_VIEWS) {
        napi_throw_type_error(env, NULL, "batchRefCounts(ctx, nViews)");
        return NULL;
    }
    CHECK(env, c, gs3d_batch_ref_counts(c, k, n), "gs3d_batch_ref_counts");
    double d[GS3D_MAX_BATCH_VIEWS];
    for (uint32_t i = 0; i < n; i++) d[i] = (double)k[i];
    return make_typed(env, napi_float64_array, d, n, 8);
}

/* batchRead(ctx, view, bufferId) -> ArrayBuffer */
static napi_value js_batch_read(napi_env env, napi_callback_info info) {
    ARGS(3);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint32_t view = 0;
    int32_t id = 0;
    if (!arg_u32(env, argv[1], &view, "batchRead(ctx, view, bufferId)") ||
        !arg_i32(env, argv[2], &id, "batchRead(ctx, view, bufferId)"))
        return NULL;
    size_t size = 0;
    CHECK(env, c, gs3d_batch_read(c, view, id, NULL, 0, &size), "gs3d_batch_read");
    void* data;
    napi_value ab;
    NAPI_CALL(env, napi_create_arraybuffer(env, size, &data, &ab));
    CHECK(env, c, gs3d_batch_read(c, view, id, data, size, &size), "gs3d_batch_read");
    return ab;
}

/* learnedInit(ctx, seed) */
static napi_value js_learned_init(napi_env env, napi_callback_info info) {
    ARGS(2);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    int64_t seed = 0;
    if (argc > 1 && !is_nullish(env, argv[1]) && napi_get_value_int64(env, argv[1], &seed) != napi_ok) {
        napi_throw_type_error(env, NULL, "learnedInit(ctx, seed?)");
        return NULL;
    }
    CHECK(env, c, gs3d_learned_init(c, (uint64_t)seed), "gs3d_learned_init");
    return NULL;
}

/* learnedTrain(ctx, epochs, lr, lrLatent) -> Float32Array (mean squared error per epoch) */
static napi_value js_learned_train(napi_env env, napi_callback_info info) {
    ARGS(4);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint32_t epochs = 0;
    double lr = 1e-3, lrz = 1e-2;
    if (!get_u32(env, argv[1], &epochs) || epochs == 0 || epochs > 100000) {
        napi_throw_type_error(env, NULL, "learnedTrain(ctx, epochs, lr, lrLatent)");
        return NULL;
    }
    if ((argc > 2 && !is_nullish(env, argv[2]) && napi_get_value_double(env, argv[2], &lr) != napi_ok) ||
        (argc > 3 && !is_nullish(env, argv[3]) && napi_get_value_double(env, argv[3], &lrz) != napi_ok) ||
        !(lr > 0.0) || !(lrz >= 0.0)) {
        napi_throw_type_error(env, NULL, "learnedTrain(ctx, epochs, lr > 0, lrLatent >= 0)");
        return NULL;
    }
    float* loss = (float*)calloc(epochs, sizeof(float));
    if (!loss) {
        napi_throw_error(env, NULL, "learnedTrain: out of memory");
        return NULL;
    }
    const int rc = gs3d_learned_train(c, epochs, (float)lr, (float)lrz, loss, epochs);
    if (rc != GS3D_OK) {
        free(loss);
        return throw_rc(env, c, rc, "gs3d_learned_train");
    }
    napi_value out = make_typed(env, napi_float32_array, loss, epochs, 4);
    free(loss);
    return out;
}

/* learnedValidate(ctx, views, minPsnrDb) -> {psnr, enabled} */
static napi_value js_learned_validate(napi_env env, napi_callback_info info) {
    ARGS(3);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    uint32_t n = 0;
    gs3d_view* v = get_views(env, argv[1], &n);
    double gate = 45.0;
    if (!v) {
        napi_throw_type_error(env, NULL, "learnedValidate(ctx, views, minPsnrDb)");
        return NULL;
    }
    if (argc > 2) napi_get_value_double(env, argv[2], &gate);
    float psnr = 0.f;
    int enabled = 0;
    const int rc = gs3d_learned_validate(c, v, n, (float)gate, &psnr, &enabled);
    free(v);
    CHECK(env, c, rc, "gs3d_learned_validate");
    napi_value r, p, e;
    napi_create_object(env, &r);
    napi_create_double(env, psnr, &p);
    napi_get_boolean(env, enabled != 0, &e);
    napi_set_named_property(env, r, "psnr", p);
    napi_set_named_property(env, r, "enabled", e);
    return r;
}

/* learnedEnable(ctx, on) */
static napi_value js_learned_enable(napi_env env, napi_callback_info info) {
    ARGS(2);
    gs3d_ctx* c = get_ctx(env, argv[0]);
    if (!c) return NULL;
    bool on = false;
    if (!arg_bool(env, argv[1], &on, "learnedEnable(ctx, on: boolean)")) return NULL;
    CHECK(env, c, gs3d_learned_enable(c, on ? 1 : 0), "gs3d_learned_enable");
    return NULL;
}

static napi_value init(napi_env env, napi_value exports) {
    napi_property_descriptor d[] = {
#define FN(js, c) {js, NULL, c, NULL, NULL, NULL, napi_enumerable, NULL}
        FN("create", js_create),
        FN("destroy", js_destroy),
        FN("lastError", js_last_error),
        FN("sync", js_sync),
        FN("deviceLimits", js_device_limits),
        FN("setSplats", js_set_splats),
        FN("setTileSize", js_set_tile_size),
        FN("getTileSize", js_get_tile_size),
        FN("getTileCount", js_get_tile_count),
        FN("rebin", js_rebin),
        FN("refCount", js_ref_count),
        FN("read", js_read),
        FN("readTileCounts", js_read_tile_counts),
        FN("stageTimes", js_stage_times),
        FN("stageTimesAvg", js_stage_times_avg),
        FN("profilerReset", js_profiler_reset),
        FN("stageLabel", js_stage_label),
        FN("bufferCreate", js_buffer_create),
        FN("bufferResize", js_buffer_resize),
        FN("bufferSize", js_buffer_size),
        FN("bufferWrite", js_buffer_write),
        FN("bufferRead", js_buffer_read),
        FN("bufferDestroy", js_buffer_destroy),
        FN("bufferList", js_buffer_list),
        FN("setKernelVariant", js_set_kernel_variant),
        FN("getKernelVariant", js_get_kernel_variant),
        FN("setProfileMask", js_set_profile_mask),
        FN("launchGeometry", js_launch_geometry),
        FN("referencePassStages", js_reference_pass_stages),
        FN("sortPasses", js_sort_passes),
        FN("setDebugMode", js_set_debug_mode),
        FN("composite", js_composite),
        FN("parsePly", js_parse_ply),
        FN("loadPly", js_load_ply),
        FN("rebinBatch", js_rebin_batch),
        FN("batchRefCounts", js_batch_ref_counts),
        FN("batchRead", js_batch_read),
        FN("learnedInit", js_learned_init),
        FN("learnedTrain", js_learned_train),
        FN("learnedValidate", js_learned_validate),
        FN("learnedEnable", js_learned_enable),
#undef FN
    };
    NAPI_CALL(env, napi_define_properties(env, exports, sizeof d / sizeof d[0], d));
    napi_value v;
    napi_create_int32(env, gs3d_abi_version(), &v);
    napi_set_named_property(env, exports, "abiVersion", v);
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
