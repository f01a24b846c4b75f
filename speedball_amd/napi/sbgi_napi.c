/*
 * sbgi_napi.c — thin N-API addon over include/sbgi.h for the Node host.
 *
 * Every export maps 1:1 onto a C-ABI call; typed arrays are passed by pointer for the
 * duration of the call (the library copies what it keeps). A negative status throws a
 * JS Error whose message is sbgi_last_error() and whose `code` is the SBGI_E_* value,
 * so the JS shim can route SBGI_E_LAYOUT / SBGI_E_CAPACITY to requestRebuild as the
 * reference routes a failed refit or an oversized build (js/gi_probes.js:3299-3302,
 * 4170, 4220).
 *
 * Built against the N-API headers of the installed Node (napi v8):
 *   gcc -shared -fPIC -I/usr/include/node sbgi_napi.c -L.. -lsbgi -Wl,-rpath,'$ORIGIN/..'
 */
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sbgi.h"

#define NAPI_CALL(env, call)                                               \
    do {                                                                   \
        if ((call) != napi_ok) {                                           \
            napi_throw_error((env), NULL, "N-API call failed: " #call);    \
            return NULL;                                                   \
        }                                                                  \
    } while (0)

static napi_value throw_status(napi_env env, int32_t rc, const char *msg) {
    napi_value err, m, code;
    char buf[512];
    snprintf(buf, sizeof(buf), "sbgi error %d: %s", rc, msg ? msg : "");
    napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &m);
    napi_create_error(env, NULL, m, &err);
    napi_create_int32(env, rc, &code);
    napi_set_named_property(env, err, "code", code);
    napi_throw(env, err);
    return NULL;
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
    size_t argc = want;
    if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
    for (size_t i = argc; i < want; i++) napi_get_undefined(env, &argv[i]);
    return 1;
}

static void *get_ptr(napi_env env, napi_value v) {
    void *p = NULL;
    napi_valuetype t;
    if (napi_typeof(env, v, &t) != napi_ok || t != napi_external) return NULL;
    napi_get_value_external(env, v, &p);
    return p;
}

/* data pointer + element count of a typed array (NULL for null/undefined) */
static void *get_typed(napi_env env, napi_value v, size_t *len, napi_typedarray_type *type) {
    napi_valuetype t;
    *len = 0;
    if (napi_typeof(env, v, &t) != napi_ok || t == napi_null || t == napi_undefined) return NULL;
    bool is_ta = false;
    napi_is_typedarray(env, v, &is_ta);
    if (!is_ta) return NULL;
    void *data = NULL;
    napi_value ab;
    size_t off = 0;
    napi_typedarray_type tt;
    napi_get_typedarray_info(env, v, &tt, len, &data, &ab, &off);
    if (type) *type = tt;
    return data;
}

static double get_num(napi_env env, napi_value obj, const char *key, double dflt) {
    bool has = false;
    napi_value v;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
    napi_get_named_property(env, obj, key, &v);
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t != napi_number) return dflt;
    double d;
    napi_get_value_double(env, v, &d);
    return d;
}

static napi_value get_prop(napi_env env, napi_value obj, const char *key) {
    napi_value v;
    napi_get_named_property(env, obj, key, &v);
    return v;
}

static napi_value make_u32(napi_env env, uint32_t x) {
    napi_value v;
    napi_create_uint32(env, x, &v);
    return v;
}

static napi_value make_f64(napi_env env, double x) {
    napi_value v;
    napi_create_double(env, x, &v);
    return v;
}

/* ---------------------------------------------------------------- field */
static napi_value js_version(napi_env env, napi_callback_info info) {
    (void)info;
    napi_value s;
    napi_create_string_utf8(env, sbgi_version(), NAPI_AUTO_LENGTH, &s);
    return s;
}

static napi_value js_field_create(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return NULL;
    sbgi_field_desc d;
    memset(&d, 0, sizeof(d));
    d.device = (int32_t)get_num(env, argv[0], "device", 0);
    d.rays_per_probe = (uint32_t)get_num(env, argv[0], "rays", 64);
    d.depth_oct_res = (uint32_t)get_num(env, argv[0], "depthOct", 6);
    d.blend_precision = (uint32_t)get_num(env, argv[0], "precision", SBGI_BLEND_FP32);
    d.traversal = (uint32_t)get_num(env, argv[0], "traversal", SBGI_TRAV_MT_COMPAT);
    d.flags = (uint32_t)get_num(env, argv[0], "flags", 0);
    d.reflections = (uint32_t)get_num(env, argv[0], "reflections", SBGI_REFLECT_OFF);
    napi_value casc = get_prop(env, argv[0], "cascades");
    uint32_t n = 0;
    napi_get_array_length(env, casc, &n);
    if (n < 1 || n > SBGI_MAX_CASCADES) return throw_status(env, SBGI_E_INVALID, "cascades must be 1..4");
    d.cascade_count = n;
    for (uint32_t i = 0; i < n; i++) {
        napi_value c;
        napi_get_element(env, casc, i, &c);
        size_t l1, l2, l3;
        napi_typedarray_type t;
        const float *gmin = get_typed(env, get_prop(env, c, "gridMin"), &l1, &t);
        const float *gsize = get_typed(env, get_prop(env, c, "gridSize"), &l2, &t);
        const uint32_t *res = get_typed(env, get_prop(env, c, "res"), &l3, &t);
        if (!gmin || !gsize || !res || l1 != 3 || l2 != 3 || l3 != 3)
            return throw_status(env, SBGI_E_INVALID, "cascade needs Float32Array gridMin/gridSize and Uint32Array res");
        memcpy(d.cascades[i].grid_min, gmin, 12);
        memcpy(d.cascades[i].grid_size, gsize, 12);
        memcpy(d.cascades[i].res, res, 12);
    }
    sbgi_field *f = NULL;
    int32_t rc = sbgi_create(&d, &f);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(NULL));
    napi_value ext;
    NAPI_CALL(env, napi_create_external(env, f, NULL, NULL, &ext));
    return ext;
}

static napi_value js_field_destroy(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return NULL;
    sbgi_destroy((sbgi_field *)get_ptr(env, argv[0]));
    return NULL;
}

/* The scene's typed arrays with their element types and lengths checked against each
 * other, so that no count taken from one array can make sbgi_set_scene read past the end
 * of another: bvhNodes Uint32Array[8 N], triIndex Uint32Array[3 T], vertexData
 * Float32Array[8 V], triMaterial Uint32Array[T], materials Float32Array, lights
 * Float32Array[17 L] (optional). Returns NULL when the scene is well formed, else the
 * reason. */
static const char *typed_of(napi_env env, napi_value o, const char *key, napi_typedarray_type want, size_t stride,
                            int required, const void **data, size_t *count) {
    size_t l = 0;
    napi_typedarray_type t = napi_int8_array;
    *data = get_typed(env, get_prop(env, o, key), &l, &t);
    *count = 0;
    if (!*data) return required ? key : NULL;
    if (t != want || l % stride != 0) return key;
    *count = l / stride;
    return NULL;
}

static const char *read_scene(napi_env env, napi_value o, sbgi_scene_view *v) {
    const void *d;
    size_t n, tris = 0, mats = 0;
    const char *bad;
    memset(v, 0, sizeof(*v));
    if ((bad = typed_of(env, o, "bvhNodes", napi_uint32_array, 8, 1, &d, &n))) return bad;
    v->bvh_nodes = d;
    v->node_count = (uint32_t)n;
    if ((bad = typed_of(env, o, "triIndex", napi_uint32_array, 3, 1, &d, &tris))) return bad;
    v->tri_index = d;
    v->tri_count = (uint32_t)tris;
    if ((bad = typed_of(env, o, "vertexData", napi_float32_array, 8, 1, &d, &n))) return bad;
    v->vertex_data = d;
    v->vertex_count = (uint32_t)n;
    if ((bad = typed_of(env, o, "triMaterial", napi_uint32_array, 1, 1, &d, &n))) return bad;
    if (n != tris) return "triMaterial (one entry per triIndex triangle)";
    v->tri_material = d;
    if ((bad = typed_of(env, o, "materials", napi_float32_array, 1, 1, &d, &mats))) return bad;
    v->materials = d;
    v->materials_len = (uint32_t)mats;
    if ((bad = typed_of(env, o, "lights", napi_float32_array, SBGI_LIGHT_STRIDE, 0, &d, &n))) return bad;
    v->lights = d;
    v->light_count = (uint32_t)n;
    v->inst_base = (uint32_t)get_num(env, o, "instBase", 0);
    v->inst_count = (uint32_t)get_num(env, o, "instCount", 0);
    v->tlas_base = (uint32_t)get_num(env, o, "tlasBase", 0);
    v->tlas_node_count = (uint32_t)get_num(env, o, "tlasNodeCount", 0);
    v->emitter_base = (uint32_t)get_num(env, o, "emitterBase", 0);
    v->emitter_count = (uint32_t)get_num(env, o, "emitterCount", 0);
    if ((uint64_t)v->inst_base + (uint64_t)v->inst_count * SBGI_MAT_STRIDE > mats ||
        (uint64_t)v->tlas_base + (uint64_t)v->tlas_node_count * SBGI_TLAS_STRIDE > mats)
        return "materials (instance or TLAS records past its end)";
    if ((uint64_t)v->emitter_base + v->emitter_count > v->light_count) return "lights (emitter range past its end)";
    return NULL;
}

static napi_value js_field_set_scene(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    sbgi_scene_view v;
    /* the scene is checked before the field, so its checks need no device */
    const char *bad = read_scene(env, argv[1], &v);
    if (bad) {
        char msg[160];
        snprintf(msg, sizeof(msg), "malformed scene: %s has the wrong type or length", bad);
        return throw_status(env, SBGI_E_INVALID, msg);
    }
    if (!f) return throw_status(env, SBGI_E_INVALID, "bad field");
    int32_t rc = sbgi_set_scene(f, &v);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    return NULL;
}

/* tick params: every sbgi_tick_params field by its camelCase name */
static void read_params(napi_env env, napi_value o, sbgi_tick_params *p) {
    sbgi_tick_params_default(p);
    p->cascade = (uint32_t)get_num(env, o, "cascade", 0);
    p->probe_offset = (uint32_t)get_num(env, o, "probeOffset", 0);
    p->updated_count = (uint32_t)get_num(env, o, "updatedCount", 0);
    p->needs_clear = (uint32_t)get_num(env, o, "needsClear", 0);
#define F(field, key) p->field = (float)get_num(env, o, key, p->field)
    F(frame_jitter, "frameJitter");
    F(hysteresis, "hysteresis");
    F(hysteresis_exponent, "hysteresisExponent");
    F(depth_sharpness, "depthSharpness");
    F(radiance_clamp, "radianceClamp");
    F(classify_strength, "classifyStrength");
    F(filter_strength, "filterStrength");
    F(filter_smooth, "filterSmooth");
    F(temp_change_sigma1, "tempChangeSigma1");
    F(temp_change_h_drop, "tempChangeHDrop");
    F(temp_clamp_sigma, "tempClampSigma");
    F(nir_gate, "nirGate");
    F(nir_gain, "nirGain");
    F(sky_intensity, "skyIntensity");
    F(dbg_trace_bias, "debugTraceBiasScale");
    F(dbg_ray_eps, "debugRayEpsScale");
    F(dbg_direct, "debugDirectScale");
    F(dbg_emissive, "debugEmissiveScale");
    F(dbg_albedo, "debugAlbedoScale");
    F(dbg_bounce, "debugBounceScale");
    F(dbg_cosine_power, "debugCosinePower");
    F(dbg_noise_h_boost, "debugTempNoiseHBoost");
    F(dbg_change_sigma0, "debugTempChangeSigma0");
    F(dbg_min_change_h, "debugTempMinChangeH");
    F(dbg_var_eps, "debugTempVarEps");
    F(dbg_var_rel, "debugTempVarRel");
    F(dbg_depth_history, "debugDepthHistoryScale");
    F(dbg_filter_k, "debugFilterKScale");
    F(dbg_filter_rel, "debugFilterRelScale");
    F(dbg_filter_eps, "debugFilterEpsScale");
    F(emitter_vis_seed, "emitterVisSeed");
    F(sky_configured, "skyConfigured");
    F(reflection_sky_fallback, "reflectionSkyFallback");
#undef F
    p->glossy_phase = (uint32_t)get_num(env, o, "glossyPhase", 0);
    size_t l;
    napi_typedarray_type t;
    const float *sh = get_typed(env, get_prop(env, o, "skySH"), &l, &t);
    if (sh && l == 27 && t == napi_float32_array) memcpy(p->sky_sh, sh, sizeof(p->sky_sh));
}

static napi_value js_field_tick(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    if (!f) return throw_status(env, SBGI_E_INVALID, "bad field");
    sbgi_tick_params p;
    read_params(env, argv[1], &p);
    uint64_t ticket = 0;
    int32_t rc = sbgi_tick(f, &p, &ticket);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    return make_f64(env, (double)ticket);
}

static napi_value js_field_wait(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    if (!get_args(env, info, 3, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    double ticket = 0, to = -1;
    napi_get_value_double(env, argv[1], &ticket);
    napi_valuetype t;
    napi_typeof(env, argv[2], &t);
    if (t == napi_number) napi_get_value_double(env, argv[2], &to);
    int32_t rc = sbgi_wait(f, (uint64_t)ticket, (int32_t)to);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    return NULL;
}

static napi_value js_field_poll(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    double ticket = 0;
    napi_get_value_double(env, argv[1], &ticket);
    int32_t rc = sbgi_poll(f, (uint64_t)ticket);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    napi_value b;
    napi_get_boolean(env, rc == 1, &b);
    return b;
}

/* fieldClassify(field, cascade): the classifyKernel pass of one cascade (sbgi_classify) */
static napi_value js_field_classify(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    uint32_t casc = 0;
    napi_get_value_uint32(env, argv[1], &casc);
    int32_t rc = sbgi_classify(f, casc);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    return NULL;
}

/* fieldSetStageTiming(f, on): per-stage events in graph-replayed ticks (stats ms*). */
static napi_value js_field_set_stage_timing(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    bool on = false;
    napi_get_value_bool(env, argv[1], &on);
    int32_t rc = sbgi_set_stage_timing(f, on ? 1 : 0);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    return NULL;
}

static napi_value js_field_read(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    if (!get_args(env, info, 3, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    int32_t id = 0;
    uint32_t casc = 0;
    napi_get_value_int32(env, argv[1], &id);
    napi_get_value_uint32(env, argv[2], &casc);
    size_t bytes = sbgi_buffer_bytes(f, id, casc);
    if (!bytes) return throw_status(env, SBGI_E_INVALID, "buffer not available");
    void *data = NULL;
    napi_value ab;
    NAPI_CALL(env, napi_create_arraybuffer(env, bytes, &data, &ab));
    int32_t rc = sbgi_read(f, id, casc, data, bytes);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    return ab;
}

static napi_value js_field_write(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    int32_t id = 0;
    uint32_t casc = 0;
    napi_get_value_int32(env, argv[1], &id);
    napi_get_value_uint32(env, argv[2], &casc);
    size_t l;
    napi_typedarray_type t;
    const void *src = get_typed(env, argv[3], &l, &t);
    if (!src || t != napi_float32_array) return throw_status(env, SBGI_E_INVALID, "expects a Float32Array");
    int32_t rc = sbgi_write(f, id, casc, src, l * 4);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    return NULL;
}

static napi_value js_field_update_ranges(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    int32_t id = 0;
    napi_get_value_int32(env, argv[1], &id);
    size_t nr, ns;
    napi_typedarray_type t;
    const uint32_t *ranges = get_typed(env, argv[2], &nr, &t);
    const void *src = get_typed(env, argv[3], &ns, &t);
    if (!ranges || !src) return throw_status(env, SBGI_E_INVALID, "ranges and source must be typed arrays");
    int32_t rc = sbgi_update_ranges(f, id, ranges, (uint32_t)(nr / 2), src);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    return NULL;
}

static napi_value js_field_refit(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    if (!get_args(env, info, 3, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    size_t nr = 0;
    napi_typedarray_type t;
    const uint32_t *ranges = get_typed(env, argv[1], &nr, &t);
    bool tl = true;
    napi_get_value_bool(env, argv[2], &tl);
    int32_t rc = sbgi_refit(f, ranges, (uint32_t)(nr / 2), tl ? 1 : 0);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    return NULL;
}

static napi_value js_field_set_transforms(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    size_t ns, nw, na;
    napi_typedarray_type t;
    const uint32_t *slots = get_typed(env, argv[1], &ns, &t);
    const double *worlds = get_typed(env, argv[2], &nw, &t);
    const uint8_t *active = get_typed(env, argv[3], &na, &t);
    if (!slots || !worlds || nw != ns * 16 || (active && na != ns))
        return throw_status(env, SBGI_E_INVALID, "slots Uint32Array, worlds Float64Array(16 n), active Uint8Array(n)");
    int32_t rc = sbgi_set_instance_transforms(f, slots, (uint32_t)ns, worlds, active);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    return NULL;
}

static napi_value js_field_stats(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return NULL;
    sbgi_field *f = get_ptr(env, argv[0]);
    sbgi_stats s;
    int32_t rc = sbgi_get_stats(f, &s);
    if (rc < 0) return throw_status(env, rc, sbgi_last_error(f));
    napi_value o;
    napi_create_object(env, &o);
    napi_set_named_property(env, o, "msRaydir", make_f64(env, s.ms_raydir));
    napi_set_named_property(env, o, "msTrace", make_f64(env, s.ms_trace));
    napi_set_named_property(env, o, "msBlend", make_f64(env, s.ms_blend));
    napi_set_named_property(env, o, "msUpload", make_f64(env, s.ms_upload));
    napi_set_named_property(env, o, "msTotal", make_f64(env, s.ms_total));
    napi_set_named_property(env, o, "rays", make_u32(env, s.rays));
    napi_set_named_property(env, o, "updated", make_u32(env, s.updated));
    napi_set_named_property(env, o, "msWalkPrimary", make_f64(env, s.ms_walk_primary));
    napi_set_named_property(env, o, "msEmit", make_f64(env, s.ms_emit));
    napi_set_named_property(env, o, "msWalkShadow", make_f64(env, s.ms_walk_shadow));
    napi_set_named_property(env, o, "msResolve", make_f64(env, s.ms_resolve));
    napi_set_named_property(env, o, "msEmitterVis", make_f64(env, s.ms_emitter_vis));
    napi_set_named_property(env, o, "tickReplayed", make_f64(env, (double)s.tick_replayed));
    napi_set_named_property(env, o, "stageTimed", make_f64(env, (double)s.stage_timed));
    return o;
}

/* ---------------------------------------------------------------- scene builder */
static napi_value js_sb_create(napi_env env, napi_callback_info info) {
    (void)info;
    sbgi_scene_builder *b = NULL;
    int32_t rc = sbgi_sb_create(&b);
    if (rc < 0) return throw_status(env, rc, "builder allocation failed");
    napi_value ext;
    NAPI_CALL(env, napi_create_external(env, b, NULL, NULL, &ext));
    return ext;
}

static napi_value js_sb_destroy(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return NULL;
    sbgi_sb_destroy((sbgi_scene_builder *)get_ptr(env, argv[0]));
    return NULL;
}

static napi_value js_sb_add_material(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return NULL;
    sbgi_scene_builder *b = get_ptr(env, argv[0]);
    size_t l;
    napi_typedarray_type t;
    const float *rec = get_typed(env, argv[1], &l, &t);
    if (!rec || l != 28) return throw_status(env, SBGI_E_INVALID, "material record must be Float32Array(28)");
    uint32_t id = 0;
    int32_t rc = sbgi_sb_add_material(b, rec, &id);
    if (rc < 0) return throw_status(env, rc, sbgi_sb_last_error(b));
    return make_u32(env, id);
}

static napi_value js_sb_add_mesh(napi_env env, napi_callback_info info) {
    napi_value argv[6];
    if (!get_args(env, info, 6, argv)) return NULL;
    sbgi_scene_builder *b = get_ptr(env, argv[0]);
    size_t lp, ln, lu, li, lm;
    napi_typedarray_type t;
    const float *pos = get_typed(env, argv[1], &lp, &t);
    const float *nrm = get_typed(env, argv[2], &ln, &t);
    const float *uv = get_typed(env, argv[3], &lu, &t);
    const uint32_t *idx = get_typed(env, argv[4], &li, &t);
    const uint32_t *tm = get_typed(env, argv[5], &lm, &t);
    if (!pos || !idx || lp % 3 || li % 3 || (nrm && ln != lp) || (uv && lu != lp / 3 * 2) || (tm && lm != li / 3))
        return throw_status(env, SBGI_E_INVALID, "mesh arrays have inconsistent sizes");
    uint32_t id = 0;
    int32_t rc = sbgi_sb_add_mesh(b, pos, (uint32_t)(lp / 3), nrm, uv, idx, (uint32_t)(li / 3), tm, &id);
    if (rc < 0) return throw_status(env, rc, sbgi_sb_last_error(b));
    return make_u32(env, id);
}

static napi_value js_sb_add_instance(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return NULL;
    sbgi_scene_builder *b = get_ptr(env, argv[0]);
    uint32_t mesh = 0;
    napi_get_value_uint32(env, argv[1], &mesh);
    size_t l;
    napi_typedarray_type t;
    const double *w = get_typed(env, argv[2], &l, &t);
    bool active = true;
    napi_get_value_bool(env, argv[3], &active);
    if (!w || l != 16) return throw_status(env, SBGI_E_INVALID, "world matrix must be Float64Array(16)");
    uint32_t id = 0;
    int32_t rc = sbgi_sb_add_instance(b, mesh, w, active ? 1 : 0, &id);
    if (rc < 0) return throw_status(env, rc, sbgi_sb_last_error(b));
    return make_u32(env, id);
}

static napi_value js_sb_add_light(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return NULL;
    sbgi_scene_builder *b = get_ptr(env, argv[0]);
    size_t l;
    napi_typedarray_type t;
    const float *rec = get_typed(env, argv[1], &l, &t);
    if (!rec || l != 17) return throw_status(env, SBGI_E_INVALID, "light record must be Float32Array(17)");
    int32_t rc = sbgi_sb_add_light(b, rec);
    if (rc < 0) return throw_status(env, rc, sbgi_sb_last_error(b));
    return NULL;
}

static napi_value copy_typed(napi_env env, const void *src, size_t n, napi_typedarray_type t, size_t esz) {
    void *data = NULL;
    napi_value ab, ta;
    napi_create_arraybuffer(env, n * esz, &data, &ab);
    if (n && src) memcpy(data, src, n * esz);
    napi_create_typedarray(env, t, n, ab, 0, &ta);
    return ta;
}

static napi_value js_sb_finalize(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return NULL;
    sbgi_scene_builder *b = get_ptr(env, argv[0]);
    sbgi_scene_view v;
    int32_t rc = sbgi_sb_finalize(b, &v);
    if (rc < 0) return throw_status(env, rc, sbgi_sb_last_error(b));
    napi_value o;
    napi_create_object(env, &o);
    napi_set_named_property(env, o, "bvhNodes", copy_typed(env, v.bvh_nodes, (size_t)v.node_count * 8, napi_uint32_array, 4));
    napi_set_named_property(env, o, "triIndex", copy_typed(env, v.tri_index, (size_t)v.tri_count * 3, napi_uint32_array, 4));
    napi_set_named_property(env, o, "vertexData", copy_typed(env, v.vertex_data, (size_t)v.vertex_count * 8, napi_float32_array, 4));
    napi_set_named_property(env, o, "triMaterial", copy_typed(env, v.tri_material, v.tri_count, napi_uint32_array, 4));
    napi_set_named_property(env, o, "materials", copy_typed(env, v.materials, v.materials_len, napi_float32_array, 4));
    napi_set_named_property(env, o, "lights", copy_typed(env, v.lights, (size_t)v.light_count * 17, napi_float32_array, 4));
    napi_set_named_property(env, o, "instBase", make_u32(env, v.inst_base));
    napi_set_named_property(env, o, "instCount", make_u32(env, v.inst_count));
    napi_set_named_property(env, o, "tlasBase", make_u32(env, v.tlas_base));
    napi_set_named_property(env, o, "tlasNodeCount", make_u32(env, v.tlas_node_count));
    napi_set_named_property(env, o, "emitterBase", make_u32(env, v.emitter_base));
    napi_set_named_property(env, o, "emitterCount", make_u32(env, v.emitter_count));
    uint32_t *order = malloc(sizeof(uint32_t) * (v.inst_count ? v.inst_count : 1));
    if (order && sbgi_sb_tlas_order(b, order, v.inst_count) == SBGI_OK)
        napi_set_named_property(env, o, "tlasOrder", copy_typed(env, order, v.inst_count, napi_uint32_array, 4));
    free(order);
    return o;
}

static napi_value js_sb_set_transform(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    if (!get_args(env, info, 3, argv)) return NULL;
    sbgi_scene_builder *b = get_ptr(env, argv[0]);
    uint32_t inst = 0;
    napi_get_value_uint32(env, argv[1], &inst);
    size_t l;
    napi_typedarray_type t;
    const double *w = get_typed(env, argv[2], &l, &t);
    if (!w || l != 16) return throw_status(env, SBGI_E_INVALID, "world matrix must be Float64Array(16)");
    uint32_t slot = 0;
    int32_t rc = sbgi_sb_set_transform(b, inst, w, &slot);
    if (rc < 0) return throw_status(env, rc, sbgi_sb_last_error(b));
    return make_u32(env, slot);
}

#define EXPORT(name, fn)                                                       \
    do {                                                                       \
        napi_value f_;                                                         \
        napi_create_function(env, name, NAPI_AUTO_LENGTH, fn, NULL, &f_);      \
        napi_set_named_property(env, exports, name, f_);                       \
    } while (0)

static napi_value init(napi_env env, napi_value exports) {
    EXPORT("version", js_version);
    EXPORT("fieldCreate", js_field_create);
    EXPORT("fieldDestroy", js_field_destroy);
    EXPORT("fieldSetScene", js_field_set_scene);
    EXPORT("fieldTick", js_field_tick);
    EXPORT("fieldClassify", js_field_classify);
    EXPORT("fieldWait", js_field_wait);
    EXPORT("fieldPoll", js_field_poll);
    EXPORT("fieldRead", js_field_read);
    EXPORT("fieldWrite", js_field_write);
    EXPORT("fieldUpdateRanges", js_field_update_ranges);
    EXPORT("fieldRefit", js_field_refit);
    EXPORT("fieldSetTransforms", js_field_set_transforms);
    EXPORT("fieldStats", js_field_stats);
    EXPORT("fieldSetStageTiming", js_field_set_stage_timing);
    EXPORT("sbCreate", js_sb_create);
    EXPORT("sbDestroy", js_sb_destroy);
    EXPORT("sbAddMaterial", js_sb_add_material);
    EXPORT("sbAddMesh", js_sb_add_mesh);
    EXPORT("sbAddInstance", js_sb_add_instance);
    EXPORT("sbAddLight", js_sb_add_light);
    EXPORT("sbFinalize", js_sb_finalize);
    EXPORT("sbSetTransform", js_sb_set_transform);
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
