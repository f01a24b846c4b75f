/*
 * sbgi.h — C ABI of the MI355X-native DDGI probe update.
 *
 * This is the boundary that replaces the reference's WebGPU submission path:
 * three.js StorageBufferAttribute uploads plus renderer.computeAsync over
 * [traceKernel, blendKernel, uploadKernel] (js/gi_probes.js:1484-1561,
 * 3812-3836), the ranged uploads of markStorageDirty/addUpdateRange
 * (js/gi_probes.js:1543-1561, 4239-4251), the CPU refit lanes
 * (js/gi_refit.js:34-430, js/spectral_scene.js:1402-1514, 1831-1974) and the
 * debug readback getArrayBufferAsync (js/gi_probes.js:4830-4841).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an int32_t status: 0 on success, a negative SBGI_E_* code on
 * failure (the message is available from sbgi_last_error). Host arrays are
 * borrowed for the duration of the call only; device memory is owned by the
 * sbgi_field handle.
 *
 * Buffers passed to sbgi_set_scene use the reference byte layouts unchanged:
 *   bvhNodes     u32[8*N]  [f32 min.xyz, f32 max.xyz, u32 miss, u32 payload]
 *                          payload 0xFFFFFFFF = interior, else
 *                          (count << 24) | pool-absolute triangle offset
 *                          (js/spectral_scene.js:1225-1234)
 *   triIndex     u32[3*T]  pool-absolute vertex ids (js/spectral_scene.js:1218-1224)
 *   vertexData   f32[8*V]  pos.xyz normal.xyz uv.xy, local space
 *   triMaterial  u32[T]    uber record index
 *   materials    f32       uber records x28 | instance records x28 at inst_base
 *                          | TLAS nodes x12 at tlas_base (js/spectral_scene.js:1308-1366)
 *   lights       f32[17*L] light records (js/spectral_scene.js:723-813)
 */
#ifndef SBGI_H
#define SBGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define SBGI_OK 0
#define SBGI_E_INVALID -1   /* bad argument (null pointer, size out of range)      */
#define SBGI_E_LAYOUT -2    /* threaded BVH layout violated; caller must rebuild   */
#define SBGI_E_CAPACITY -3  /* scene exceeds a 24-bit index ceiling or MAX_LIGHTS  */
#define SBGI_E_DEVICE -4    /* HIP runtime error                                   */
#define SBGI_E_STATE -5     /* call out of order (no scene, tick in flight ...)    */
#define SBGI_E_TIMEOUT -6   /* sbgi_wait timed out                                 */
#define SBGI_E_NOMEM -7     /* device or host allocation failed                    */

/* ---- reference constants (js/gi_probes.js:53-56, 97, 132; spectral_traverse.js:34-35) */
#define SBGI_OCT_RES 6
#define SBGI_BORDER 1
#define SBGI_TILE 8
#define SBGI_RAYS_MIN 32
#define SBGI_RAYS_MAX 256
#define SBGI_MAX_LIGHTS 64
#define SBGI_LIGHT_STRIDE 17
#define SBGI_MAT_STRIDE 28
#define SBGI_TLAS_STRIDE 12
#define SBGI_MAX_CASCADES 4

/* blend arithmetic */
#define SBGI_BLEND_FP32 0   /* plain FP32 vector ALU                                   */
#define SBGI_BLEND_BF16 1   /* BF16 MFMA inputs, FP32 accumulate                         */
#define SBGI_BLEND_BF16X3 2 /* BF16 hi/lo split on both operands (3 MFMAs), FP32 acc.  */
#define SBGI_BLEND_FP16 3   /* FP16 MFMA inputs, FP32 accumulate                         */

/* closest-hit traversal */
#define SBGI_TRAV_MT_COMPAT 0   /* stackless left-first threaded walk + Moller-Trumbore,
                                   the reference's visit order and epsilons */
#define SBGI_TRAV_MT_ORDERED 1  /* same stackless walk over a binned-SAH TLAS and near-first
                                   octant copies of the BLAS pool; same triangle test, so
                                   results differ only in which of two exactly equal-t
                                   hits wins */
#define SBGI_TRAV_WATERTIGHT 2  /* the MT_ORDERED walk with the watertight triangle test
                                   (Woop, Benthin and Wald 2013) on the stored vertices in
                                   place of Moller-Trumbore: a ray through an edge or vertex
                                   shared by two triangles hits one of them; hits differ from
                                   MT's only at edges and near ties (no reference
                                   counterpart: the reference has Moller-Trumbore only,
                                   js/spectral_traverse.js:388-409) */

/* local reflection lobes from the same rays (reflectionQuality, REFLECTION_QUALITY_TIERS
   js/gi_probes.js:60-73): the rough power-8 lobe in the irradiance tiles, and the glossy
   power-64 cache at glossyOct x glossyOct (+ border) per probe */
#define SBGI_REFLECT_OFF 0
#define SBGI_REFLECT_ROUGH 1    /* rough lobe only                                      */
#define SBGI_REFLECT_HIGH 2     /* + glossy 8x8, every texel solved on alternate updates */
#define SBGI_REFLECT_ULTRA 3    /* + glossy 16x16, every texel solved every update       */

/* buffer ids for sbgi_read / sbgi_write / sbgi_update_ranges */
#define SBGI_BUF_BVH_NODES 0    /* u32[8*N]                 (update_ranges)           */
#define SBGI_BUF_TRI_INDEX 1    /* u32[3*T]                 (update_ranges, read)     */
#define SBGI_BUF_VERTEX_DATA 2  /* f32[8*V]                 (update_ranges, read)     */
#define SBGI_BUF_TRI_MATERIAL 3 /* u32[T]                   (update_ranges, read)     */
#define SBGI_BUF_MATERIALS 4    /* f32 materials|inst|tlas  (update_ranges, read)     */
#define SBGI_BUF_LIGHTS 5       /* f32[17*L]                (update_ranges, read)     */
#define SBGI_BUF_IRR_STATE 10   /* f32[4*P*64]    per cascade (read, write)            */
#define SBGI_BUF_DEPTH_STATE 11 /* f32[2*P*TD*TD] per cascade (read, write)            */
#define SBGI_BUF_META 12        /* f32[4*P]       per cascade (read, write)            */
#define SBGI_BUF_IRR_ATLAS 13   /* u16 RGBA16F [(rx*8) x (ry*rz*8)]      (read)        */
#define SBGI_BUF_DEPTH_ATLAS 14 /* u16 RGBA16F [(rx*TD) x (ry*rz*TD)]    (read)        */
#define SBGI_BUF_RAY_DATA 15    /* f32[4*updated*R] reference order (slot*R+k)  (read) */
#define SBGI_BUF_HIT_IDS 16     /* i32[2*updated*R] (tri, inst) per ray, same order    */
#define SBGI_BUF_NODES_REFIT 17 /* bvhNodes after a device refit (read)               */
#define SBGI_BUF_TRI_PACKED 18  /* f32[12*T] packed triangle records (read, debug)    */
#define SBGI_BUF_INST_BOUNDS 19 /* f32[6*I] instance world AABBs per TLAS slot (read)  */
#define SBGI_BUF_TLAS_SAH 20    /* u32[8*(2I-1)] MT_ORDERED SAH TLAS, walk format {min.xyz,
                                   miss},{max.xyz, payload} (read)                     */
#define SBGI_BUF_EMITTER_VIS 21 /* f32[16*P] per cascade: retained emitter visibility,
                                   probe-major, 16 = SBGI_MAX_EMITTERS (read, write)  */
#define SBGI_BUF_LIGHT_GRID 22  /* u32[17*cells] clustered light lists of cascade 0 (read):
                                   per cell {count, light ids[16]}, count 17 = overflow */
#define SBGI_BUF_ROUGH_STATE 23 /* f32[4*P*64] rough lobe {rgb x coverage, coverage} per
                                   irradiance texel, per cascade (read, write)          */
#define SBGI_BUF_ROUGH_ATLAS 24 /* u16 RGBA16F, the irradiance atlas's layout (read)     */
#define SBGI_BUF_GLOSSY_STATE 25 /* f32[4*P*G*G] glossy numerator, G = glossy tile      */
#define SBGI_BUF_GLOSSY_WEIGHT 26 /* f32[P*G*G] glossy support                           */
#define SBGI_BUF_GLOSSY_ATLAS 27 /* u16 RGBA16F [(tx*G) x (ty*G)], tx = ceil(sqrt(P)),
                                   ty = ceil(P / tx), probe p at (p % tx, p / tx) (read) */
#define SBGI_BUF_TLAS_OCT 28    /* u32[8*8*(2I-1)] the SAH TLAS's near-first copies, one per
                                   ray direction octant (bit 0: -x, 1: -y, 2: -z), that the
                                   closest-hit walk reads; MT_ORDERED / WATERTIGHT scenes of
                                   at most 512 instances, else 0 bytes (read)           */
#define SBGI_BUF_SHADOW_VIS 29  /* u64[updated*R] the last solve's light visibility per
                                   ray, same order as RAY_DATA: bit l set when the ray's
                                   hit passed light l's NEE gate and the shadow ray toward
                                   l is unblocked (read)                               */
/* The shadow walk's light-space candidate lists (no reference counterpart; 0 bytes when
   the field walks shadow rays over the TLAS). PROJ: f32[16*L] per light {kind (0 point /
   spot, 1 directional, 2 none), near radius, range cut, 0}, then {position, 0} or {e1,
   u0}, {e2, v0}, {su, sv, margin, 0}; CELLS: u32[2*L*1538] {first, end} per cell record
   (6 x 16 x 16 cube-map cells, or a 32 x 32 grid for directional lights, then {0, I} and
   {I, I+1}); LIST: u16[capacity] instance slots: the identity list, then entry I (the
   never-hit box of cells no instance reaches), then the cells' lists, each nearest its
   light first (ties by slot) (read). */
#define SBGI_BUF_SHADOW_PROJ 30
#define SBGI_BUF_SHADOW_CELLS 31
#define SBGI_BUF_SHADOW_LIST 32

/* ---- field description ---------------------------------------------------- */
typedef struct sbgi_cascade_desc {
    float grid_min[3];  /* world-space grid origin (U.gridMin)                       */
    float grid_size[3]; /* world-space extent (U.gridSize)                          */
    uint32_t res[3];    /* probes per axis (U.resX/Y/Z)                             */
    uint32_t pad;
} sbgi_cascade_desc;

typedef struct sbgi_field_desc {
    int32_t device;              /* HIP device ordinal                               */
    uint32_t rays_per_probe;     /* R, multiple of 16 in [32, 256]                   */
    uint32_t cascade_count;      /* 1..SBGI_MAX_CASCADES                              */
    uint32_t depth_oct_res;      /* 6: depth shares the 8x8 tile (reference);
                                    14: 16x16 depth tile                             */
    uint32_t blend_precision;    /* SBGI_BLEND_*                                     */
    uint32_t traversal;          /* SBGI_TRAV_*                                      */
    uint32_t flags;              /* SBGI_FIELD_* bits                                */
    uint32_t reflections;        /* SBGI_REFLECT_* (reflectionQuality)               */
    sbgi_cascade_desc cascades[SBGI_MAX_CASCADES];
} sbgi_field_desc;

#define SBGI_FIELD_KEEP_HIT_IDS 1u /* store per-ray (tri, inst) for parity checks */
#define SBGI_FIELD_CLUSTERED 2u    /* clustered NEE: per-cell light lists over cascade 0's
                                      cells (clusteredLighting, js/gi_probes.js:1097,
                                      1888-1938, 2064-2127); needs cascade 0 res >= 2 on
                                      every axis. Same radiance as the all-lights loop.  */
#define SBGI_FIELD_NO_GRAPH 4u     /* every tick launches its kernels eagerly instead of
                                      replaying the tick's captured hipGraph             */
#define SBGI_FIELD_NO_SHADOW_LISTS 8u /* walk shadow rays over the TLAS instead of the
                                      light-space candidate lists (no reference
                                      counterpart: the lists are this build's
                                      acceleration of traverseAny; same answers)         */

/* Emitter lights (type 3 records, js/spectral_scene.js:815-875) injected as virtual rays
   (js/gi_probes.js:133-134, 2216-2310): at most this many per scene. */
#define SBGI_MAX_EMITTERS 16
#define SBGI_EMITTER_RETENTION 0.8f /* GI_EMITTER_VIS_RETENTION                       */
#define SBGI_LIGHTS_PER_CELL 16     /* GI_LIGHTS_PER_CELL                             */

/* ---- scene view (reference layouts, borrowed host pointers) ---------------- */
typedef struct sbgi_scene_view {
    const uint32_t *bvh_nodes;
    uint32_t node_count;
    uint32_t tri_count;
    const uint32_t *tri_index;
    const float *vertex_data;
    uint32_t vertex_count;
    uint32_t materials_len;      /* floats in materials                              */
    const uint32_t *tri_material;
    const float *materials;
    uint32_t inst_base;          /* float index of instance record 0                 */
    uint32_t inst_count;
    uint32_t tlas_base;          /* float index of TLAS node 0                       */
    uint32_t tlas_node_count;
    const float *lights;         /* 17 floats per light                              */
    uint32_t light_count;        /* U.lightCount: analytic lights and emitters       */
    uint32_t emitter_base;       /* U.emitterBase: first emitter record (type 3)     */
    uint32_t emitter_count;      /* U.emitterCount <= SBGI_MAX_EMITTERS              */
    uint32_t pad;
} sbgi_scene_view;

/* ---- per-solve parameters: the uniform set U of js/gi_probes.js:1198-1255 ---- */
typedef struct sbgi_tick_params {
    uint32_t cascade;            /* which cascade to solve                           */
    uint32_t probe_offset;       /* U.probeOffset (round-robin cursor)               */
    uint32_t updated_count;      /* U.updatedCount; 0 means the whole cascade        */
    uint32_t needs_clear;        /* clear atlases first (C.needsClear)               */
    float frame_jitter;          /* U.frameJitter                                    */
    float hysteresis;            /* U.hysteresis                                     */
    float hysteresis_exponent;   /* U.hysteresisExponent                             */
    float depth_sharpness;       /* U.depthSharpness (7)                             */
    float radiance_clamp;        /* U.radianceClamp (8)                              */
    float classify_strength;     /* U.classifyStrength (0)                           */
    float filter_strength;       /* U.filterStrength (1)                             */
    float filter_smooth;         /* U.filterSmooth (0.5)                             */
    float temp_change_sigma1;    /* U.tempChangeSigma1 (2.5)                         */
    float temp_change_h_drop;    /* U.tempChangeHDrop (0.30)                         */
    float temp_clamp_sigma;      /* U.tempClampSigma (6)                             */
    float nir_gate;              /* U.nirGate (0)                                    */
    float nir_gain;              /* U.nirGain (1)                                    */
    float sky_intensity;         /* U.skyIntensity (1)                               */
    float sky_sh[27];            /* U.skySH[0..8] as xyz triples (zero = black sky)  */
    float dbg_trace_bias;        /* U.debugTraceBiasScale (1)                        */
    float dbg_ray_eps;           /* U.debugRayEpsScale (1)                           */
    float dbg_direct;            /* U.debugDirectScale (1)                           */
    float dbg_emissive;          /* U.debugEmissiveScale (1)                         */
    float dbg_albedo;            /* U.debugAlbedoScale (1)                           */
    float dbg_bounce;            /* U.debugBounceScale (1)                           */
    float dbg_cosine_power;      /* U.debugCosinePower (1)                           */
    float dbg_noise_h_boost;     /* U.debugTempNoiseHBoost (0.25)                    */
    float dbg_change_sigma0;     /* U.debugTempChangeSigma0 (0.75)                   */
    float dbg_min_change_h;      /* U.debugTempMinChangeH (0.55)                     */
    float dbg_var_eps;           /* U.debugTempVarEps (2.5e-5)                       */
    float dbg_var_rel;           /* U.debugTempVarRel (2.5e-3)                       */
    float dbg_depth_history;     /* U.debugDepthHistoryScale (1)                     */
    float dbg_filter_k;          /* U.debugFilterKScale (1)                          */
    float dbg_filter_rel;        /* U.debugFilterRelScale (1)                        */
    float dbg_filter_eps;        /* U.debugFilterEpsScale (1)                        */
    float emitter_vis_seed;      /* U.emitterVisSeed (0): emitter sample jitter      */
    float sky_configured;        /* U.skyConfigured (0): a sky SH was supplied        */
    float reflection_sky_fallback; /* U.reflectionSkyFallback (0): misses count as
                                    reflection coverage when the sky is on            */
    uint32_t glossy_phase;       /* U.glossyPhase: texels solved this update (HIGH)  */
    uint32_t pad;
} sbgi_tick_params;

/* per-stage device time of the last completed tick, from HIP events */
typedef struct sbgi_stats {
    float ms_raydir;
    float ms_trace;
    float ms_blend;
    float ms_upload;
    float ms_total;
    uint32_t rays;               /* primary rays traced                               */
    uint32_t updated;            /* probes updated                                    */
    uint64_t shadow_rays;        /* any-hit rays launched (when counting is enabled)  */
    uint64_t blas_nodes;         /* BLAS node visits (counting builds)                */
    uint64_t tlas_nodes;         /* TLAS node visits                                  */
    uint64_t tri_tests;          /* triangle tests                                    */
    uint64_t inst_visits;        /* instance entries                                  */
    /* trace sub-stages of the last tick (ms): closest-hit walk, shadow-ray emission,
       shadow any-hit walk, shading; their sum is ms_trace minus the clears            */
    float ms_walk_primary;
    float ms_emit;
    float ms_walk_shadow;
    float ms_resolve;
    /* the shadow walk's share of the counts above (counting enabled)                 */
    uint64_t shadow_blas_nodes;
    uint64_t shadow_tlas_nodes;
    uint64_t shadow_tri_tests;
    uint64_t shadow_inst_visits;
    /* wave steps of the walks' NODE, TRI and INST phases and of lane refills (counting
       enabled): lane visits / (64 x wave steps) is the SIMD utilisation of a phase     */
    uint64_t walk_waves[4];
    uint64_t shadow_waves[4];
    float ms_emitter_vis;        /* emitter visibility rays (scenes with emitters)    */
    uint32_t tick_replayed;      /* 1 when the last tick replayed its captured hipGraph */
    uint32_t stage_timed;        /* 1 when the ms_* stage times above belong to the last
                                    tick: an eager tick, or a graph tick with stage timing
                                    on (sbgi_set_stage_timing); else they are 0          */
} sbgi_stats;

/* Wave schedule of the persistent walks (no reference counterpart: it is how this build
 * schedules the reference's walk on 64-lane waves). A triangle / instance phase runs when
 * the lanes waiting in it are >= 1/tri_ratio (1/inst_ratio) of the lanes still walking
 * nodes; finished lanes are refilled in batches of >= refill_min; rays are fetched
 * fetch_chunk at a time. Results do not depend on the schedule. shadow_lists changes which
 * boxes cull a shadow ray (its list's instance boxes instead of TLAS nodes), so only a ray
 * that grazes one of those boxes could get another answer; the GPU tests find none. */
typedef struct sbgi_walk_tuning {
    uint32_t tri_ratio;   /* 1..64, default 2 */
    uint32_t inst_ratio;  /* 1..64, default 4 */
    uint32_t refill_min;  /* 1..64, default 16 */
    uint32_t fetch_chunk; /* 1..1024, default 64 */
    uint32_t shadow_lists; /* 0 or 1: single-field ticks walk shadow rays over the
                              light-space candidate lists (SBGI_BUF_SHADOW_LIST) instead
                              of the TLAS; ignored when the field has no lists. Default
                              (set by sbgi_set_scene): 1. Batches walk the TLAS.       */
    uint32_t shadow_copies; /* 8 or 1: BLAS walk copies of shadow rays in MT_ORDERED /
                              WATERTIGHT fields: 8, the near-first copy of the ray's
                              direction octant (as closest-hit rays); 1, one copy for
                              every shadow ray, a smaller working set. Default (set by
                              sbgi_set_scene): 8 when the walk data fits 16 MiB, else 1.
                              Used by ticks and batches alike.                        */
} sbgi_walk_tuning;

typedef struct sbgi_field sbgi_field;

/* Fill *p with the reference defaults (js/gi_probes.js:1198-1255, gated jitter mode). */
void sbgi_tick_params_default(sbgi_tick_params *p);

/* Library version string. */
const char *sbgi_version(void);

/* createProbeField + buildKernels resource allocation (js/gi_probes.js:945-1046,
 * 1732-1867): allocates per-cascade state, atlases and ray scratch. */
int32_t sbgi_create(const sbgi_field_desc *desc, sbgi_field **out);

/* prepareSharedResources (js/gi_probes.js:1484-1561, 1594): copies the scene in the
 * reference layouts to device memory and derives the packed traversal layouts.
 * Fails with SBGI_E_CAPACITY for >=2^24 triangles/nodes/instances or >64 lights
 * (js/spectral_scene.js:1186, 1196-1203; js/gi_probes.js:132). */
int32_t sbgi_set_scene(sbgi_field *f, const sbgi_scene_view *scene);

/* copySceneStorageRanges / markStorageDirty (js/gi_probes.js:1543-1561, 4239-4251):
 * ranges holds n_ranges (start, count) pairs in ELEMENTS of buffer buf_id; the data
 * is taken from the same offsets of the full host array src. Vertex and light ranges
 * are enqueued without waiting for the device; structural buffers are validated first. */
int32_t sbgi_update_ranges(sbgi_field *f, int32_t buf_id, const uint32_t *ranges,
                           uint32_t n_ranges, const void *src);

/* refitFlatBlasRange (js/gi_refit.js:158-165) on the device for each [root, end)
 * pair, then, when refit_tlas != 0, the instance world bounds from the refit roots and
 * the TLAS: refitFlatTlasRange (js/gi_refit.js:245-292) in SBGI_TRAV_MT_COMPAT mode, a
 * fresh SAH build in SBGI_TRAV_MT_ORDERED mode. Returns SBGI_E_LAYOUT when the threaded
 * layout is violated, as the reference's refit returns false. Enqueued on the field's
 * stream; nothing is copied back. */
int32_t sbgi_refit(sbgi_field *f, const uint32_t *blas_ranges, uint32_t n_ranges,
                   int32_t refit_tlas);

/* updateTransforms (js/spectral_scene.js:1402-1514): rewrite the instance records of
 * the given TLAS slots from column-major world matrices (16 doubles each), recompute
 * every slot's world AABB from its BLAS root bounds on the device and update the TLAS
 * (refit in SBGI_TRAV_MT_COMPAT mode, SAH rebuild in SBGI_TRAV_MT_ORDERED mode). One
 * staged upload; enqueued on the field's stream. */
int32_t sbgi_set_instance_transforms(sbgi_field *f, const uint32_t *slots, uint32_t n,
                                     const double *world_matrices, const uint8_t *active);

/* Device skinning (no reference counterpart: the reference receives skinned positions
 * from its host and re-gathers them into vertexData in updateDeforms,
 * js/spectral_scene.js:1831-1927, then refits, js/gi_refit.js:158-165). A skin set is a
 * pooled vertex range and the BLAS over it; its rest pose is that range of the
 * vertexData last provided by the host (sbgi_set_scene / sbgi_update_ranges). Ranges
 * must be disjoint; joints index one bone palette of n_bones. n_sets = 0 clears. */
typedef struct sbgi_skin_set {
    uint32_t vertex_start, vertex_count; /* pooled vertex rows                          */
    uint32_t blas_root, blas_end;        /* threaded BLAS node range over those vertices */
    const uint16_t *joints;              /* 4 per vertex                                */
    const float *weights;                /* 4 per vertex                                */
} sbgi_skin_set;
int32_t sbgi_set_skins(sbgi_field *f, const sbgi_skin_set *sets, uint32_t n_sets, uint32_t n_bones);

/* One skinning frame: uploads the palette (n_bones 3x4 row-major matrices, 12 floats
 * each), writes p' = sum_q w_q (B_q p) and n' = normalize(sum_q w_q B_q n) into
 * vertexData for every registered vertex, refits the registered BLASes, updates their
 * walk copies and packed triangles, and, when update_tlas != 0, the instance bounds and
 * TLAS as sbgi_refit does. Enqueued on the field's stream. */
int32_t sbgi_skin(sbgi_field *f, const float *bones, uint32_t n_bones, int32_t update_tlas);

/* Rebuild the BLASes of the given [root, end) node ranges on the device from the current
 * vertex positions (no reference counterpart: the reference rebuilds with three-mesh-bvh
 * on the CPU, js/spectral_scene.js:898-966, and otherwise refits, js/gi_refit.js). For a
 * deformation that a refit would leave with swollen boxes. Same node count N (odd) and
 * pool range; L = (N + 1) / 2 leaves over Morton-ordered triangles; the ranges' triangles
 * are reordered (tri_index, tri_material) within their span, so triangle ids of those
 * BLASes change. Walk copies, packed triangles, instance bounds and the TLAS follow; the
 * new structure is copied into the library's host mirrors before the call returns. */
int32_t sbgi_rebuild_blas(sbgi_field *f, const uint32_t *blas_ranges, uint32_t n_ranges);

/* Device time (ms) of the last scene edits: skinning (or 0), BLAS refit or rebuild with
 * walk-copy and triangle updates, and instance records + bounds + TLAS. Waits for them. */
int32_t sbgi_edit_timing(sbgi_field *f, float *ms_skin, float *ms_blas_refit, float *ms_tlas);

/* One solve of one cascade: trace -> blend -> upload, the ordered pass of
 * renderer.computeAsync(gpu.solveKernels) (js/gi_probes.js:3729-3836). Enqueues on the
 * field's stream and returns; *ticket identifies the submission for sbgi_wait/poll. The
 * solve is one hipGraph per frame: a tick whose kernel arguments (scene layout, cascade,
 * parameters) equal the last captured tick's replays its graph, another tick is captured
 * and launched; needs_clear ticks and SBGI_FIELD_NO_GRAPH fields launch eagerly. */
int32_t sbgi_tick(sbgi_field *f, const sbgi_tick_params *p, uint64_t *ticket);

/* Block until the ticket has completed (timeout_ms < 0 waits forever). */
int32_t sbgi_wait(sbgi_field *f, uint64_t ticket, int32_t timeout_ms);

/* 1 when the ticket has completed, 0 when pending, negative on error. */
int32_t sbgi_poll(sbgi_field *f, uint64_t ticket);

/* _debugRead (js/gi_probes.js:4830-4841): synchronous device->host copy of a buffer.
 * cascade selects the per-cascade buffers. bytes must equal the buffer size. */
int32_t sbgi_read(sbgi_field *f, int32_t buf_id, uint32_t cascade, void *dst, size_t bytes);

/* Restore per-cascade state (irr/depth/meta) from host memory (test fixtures). */
int32_t sbgi_write(sbgi_field *f, int32_t buf_id, uint32_t cascade, const void *src,
                   size_t bytes);

/* Byte size of a buffer as sbgi_read expects it (0 when not allocated). */
size_t sbgi_buffer_bytes(const sbgi_field *f, int32_t buf_id, uint32_t cascade);

/* Device time of each stage of the last waited tick. */
int32_t sbgi_get_stats(sbgi_field *f, sbgi_stats *out);

/* Probe classification (classifyKernel, js/gi_probes.js:2850-2899; the tick dispatches it
 * at :3803-3806 when a cascade needs it and classifyStrength > 0): 32 fixed full-sphere
 * closest-hit rays per probe of the cascade from its grid position; a probe that sees back
 * faces in more than a quarter of its hits is buried (state 0, else 1), and one whose
 * nearest hit is a back face gets the relocation offset along that ray past the surface,
 * clamped to 0.45 x the smallest cell side. Writes every probe's meta {state, offset.xyz}
 * (SBGI_BUF_META), which the trace, the blend's emitter rays and the receiver use scaled
 * by classify_strength. Call it when the reference would set needsClassify: after a scene
 * change, and when classify_strength becomes positive. Stream-ordered with the ticks. */
int32_t sbgi_classify(sbgi_field *f, uint32_t cascade);

/* Counting mode (no reference counterpart; the measurement hook for the algorithmic
 * byte accounting): while on, ticks run an instrumented trace that accumulates node
 * visits, triangle tests, instance entries and shadow rays into sbgi_stats. Turning it
 * on or off resets the counters. */
int32_t sbgi_set_counting(sbgi_field *f, int32_t on);

/* Stage timing (no reference counterpart): while on, a tick replayed from its hipGraph
 * records the HIP events of its stages (ray directions, the two walks, emission, shading,
 * blend, upload), which sbgi_stats reports as ms_*. Off (the default), the tick's graph
 * is one captured piece without event nodes, which saves ~0.1 ms per tick on MI355X;
 * eager ticks record the events either way. Changing it recaptures the next tick. */
int32_t sbgi_set_stage_timing(sbgi_field *f, int32_t on);

/* Ray queries against the resident scene: the closest-hit / any-hit walks that the
 * trace uses (traverseClosest / traverseAny, js/spectral_traverse.js:441-474, 539-569),
 * for arbitrary rays. rays8 holds n records {ro.x, ro.y, ro.z, tmax, rd.x, rd.y, rd.z,
 * mode} with mode 0 = closest hit, 1 = any hit within (1e-3, tmax), 2 = closest hit
 * reported as blocked when it lies before tmax. out4 receives
 * {t, tri (int32 bits), inst (int32 bits), blocked} per ray. Synchronous. */
int32_t sbgi_query_rays(sbgi_field *f, const float *rays8, uint32_t n, float *out4);

/* sbgi_query_rays without device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory once running (per-frame query loops, hosts that
 * forbid allocation after start-up). Same rays, modes and results as sbgi_query_rays; the
 * staging and the walk's counters live in a library-owned static device array of 776 KiB
 * per GPU (held until the process ends), so n is limited to SBGI_NOALLOC_MAX_RAYS
 * (SBGI_E_CAPACITY above it). Every call on one GPU shares that array: calls that overlap
 * on one GPU (from several threads or fields) race. Synchronous. */
#define SBGI_NOALLOC_MAX_RAYS 16384
int32_t sbgi_query_rays_noalloc(sbgi_field *f, const float *rays8, uint32_t n, float *out4);

/* ---- receiver sampling: the probe field at N points ------------------------- */
/* The receiver's irradiance gather (GiProbeNode.sampleIrradiance / _sampleCascade /
 * _tapEW, js/gi_probes.js:480-614, 825-847) evaluated for a batch of points instead of
 * per fragment: P = X + normalize(Nvis) * normal_bias; per cascade the 8 probes of P's
 * cell with trilinear weights (+1e-4), Chebyshev visibility of the depth moments and the
 * wrap weight (both scaled in by cheby_strength) and the probe state (classify_strength),
 * irradiance fetched in the direction Ndir; finer cascades blended in over border_band
 * of their box (cascade c+1 is taken as finer than cascade c). */
typedef struct sbgi_sample_params {
    float normal_bias;           /* normalBias[0] x sampleBiasScale (0.04)            */
    float cheby_strength;        /* chebyStrength (0: pure trilinear)                 */
    float classify_strength;     /* classifyStrength (0)                              */
    float border_band;           /* borderBand (0.15)                                 */
    float cheby_bias[SBGI_MAX_CASCADES]; /* chebyBias per cascade (0)                 */
    uint32_t cascades;           /* cascades folded, 0 = all of the field's           */
    uint32_t pad;
} sbgi_sample_params;

void sbgi_sample_params_default(sbgi_sample_params *p);

/* pos, nvis: n x 3 floats (world position, visibility normal); ndir: n x 3 fetch
 * normals or NULL (= nvis); out: n x 3 irradiance. Host pointers; synchronous, after
 * the field's queued solves. ms (optional) receives the kernel's device time. */
int32_t sbgi_sample_irradiance(sbgi_field *f, const sbgi_sample_params *p, const float *pos, const float *nvis,
                               const float *ndir, uint32_t n, float *out, float *ms);

/* ---- differentiable mode ---------------------------------------------------------
 * No counterpart in the reference (its probe update is not differentiable); the build
 * adds it for gradient-based callers (inverse rendering, light/albedo fitting).
 *
 * The differentiable solve is two first-updates of cascade 0 from cleared state, so
 * each update's irradiance is the cosine-weighted mean of its rays with no history
 * term; the second update's bounce reads the first update's unrounded, unfiltered FP32
 * atlas. Blend FP32, filter off. Its output is the second update's interior irradiance
 * irr_out[probe][36][3]. Parameters: per-instance-triangle albedo (replacing the
 * material colour in kd) albedo[n][3] with n from sbgi_diff_param_count, ordered by
 * TLAS slot, then pool triangle within the slot's BLAS (sbgi_diff_albedo_map gives each
 * slot's (offset, first pool triangle)); and one scale per light on its colour x
 * intensity (light_scale[lights], NULL = all 1). */
int32_t sbgi_diff_param_count(sbgi_field *f, uint32_t *albedo_entries, uint32_t *lights);
int32_t sbgi_diff_albedo_map(sbgi_field *f, int32_t *map2, uint32_t n_slots);
int32_t sbgi_diff_forward(sbgi_field *f, const sbgi_tick_params *p, const float *albedo, const float *light_scale,
                          float *irr_out);
/* Adjoint of the last sbgi_diff_forward (hit sets frozen from it; any sbgi_tick in
 * between invalidates them: SBGI_E_STATE): gradients of sum(grad_irr * irr_out) with
 * respect to albedo (n x 3) and the light scales (lights; may be NULL). */
int32_t sbgi_diff_adjoint(sbgi_field *f, const float *grad_irr, float *grad_albedo, float *grad_light_scale);
/* Device time (HIP events) of the last diff forward (both updates, before the host
 * copy) and of the last adjoint's kernels. */
int32_t sbgi_diff_timing(sbgi_field *f, float *ms_forward, float *ms_adjoint);

/* Walk schedule of this field (see sbgi_walk_tuning). */
int32_t sbgi_get_walk_tuning(sbgi_field *f, sbgi_walk_tuning *out);
int32_t sbgi_set_walk_tuning(sbgi_field *f, const sbgi_walk_tuning *t);
/* Autotuner: times the trace of cascade p->cascade (probe range of p) under candidate
 * schedules by coordinate descent over the six parameters, keeps the fastest in the
 * field and returns it with its trace time. Traces into the ray scratch only: the
 * temporal state and atlases are not touched. */
int32_t sbgi_autotune(sbgi_field *f, const sbgi_tick_params *p, sbgi_walk_tuning *best, float *best_ms);

/* ---- batched solves ---------------------------------------------------------
 * No reference counterpart: the reference updates one field per ordered computeAsync
 * pass (js/gi_probes.js:3836). A batch groups fields of uneven size (own scenes,
 * cascades, ray counts) on one device; a frame updates every cascade of every field with
 * ONE persistent closest-hit walk and ONE persistent shadow walk over all (field, cascade)
 * segments (a device-side work queue), per-segment emission, shading, blend and upload.
 * A frame whose kernel arguments equal the last captured frame's replays one hipGraph;
 * the first frame, frames with needs_clear and frames after parameter changes or scene
 * edits that moved a buffer run eagerly and re-capture. While in a batch, fields are
 * updated only through it (edits on their own streams are ordered before / after frames). */
typedef struct sbgi_batch sbgi_batch;
int32_t sbgi_batch_create(sbgi_field *const *fields, uint32_t n, sbgi_batch **out);
/* params[i] for field i: cascade, probe_offset and updated_count are ignored (full
 * updates); needs_clear clears field i's atlases and starts its blend from scratch. */
int32_t sbgi_batch_tick(sbgi_batch *b, const sbgi_tick_params *params, uint64_t *ticket);
int32_t sbgi_batch_wait(sbgi_batch *b);
/* Device time of the last frame (ms) and whether it replayed the captured graph. */
int32_t sbgi_batch_timing(sbgi_batch *b, float *ms, int32_t *replayed);
/* Threads per block of the batched walks: 256 or 512. */
int32_t sbgi_batch_set_walk_block(sbgi_batch *b, uint32_t block);
/* Resident waves per CU of the batched walks: 0 (as many as registers and LDS allow) or
 * 4..64. Fewer waves keep a larger share of L2 per wave on scenes whose nodes do not fit. */
int32_t sbgi_batch_set_walk_waves(sbgi_batch *b, uint32_t waves_per_cu);
int32_t sbgi_batch_get_walk_waves(const sbgi_batch *b, uint32_t *waves_per_cu);
/* Autotuner: coordinate descent over the walk schedule (shared by the batch's fields), the
 * walk block size and the resident waves, timing eager frames; keeps and returns the
 * fastest (the waves through sbgi_batch_get_walk_waves). Each field keeps its own
 * shadow_lists choice for its own ticks; *best reports 0, the batch's TLAS walk. */
int32_t sbgi_batch_autotune(sbgi_batch *b, const sbgi_tick_params *params, sbgi_walk_tuning *best,
                            uint32_t *best_block, float *best_ms);
const char *sbgi_batch_last_error(const sbgi_batch *b);
int32_t sbgi_batch_destroy(sbgi_batch *b);

/* Last error message of this field (or of the last failed sbgi_create when f is NULL). */
const char *sbgi_last_error(const sbgi_field *f);

/* Release all device memory. Idempotent on NULL. */
int32_t sbgi_destroy(sbgi_field *f);

/* ---- scene builder: buildSpectralScene's BLAS/pool/TLAS assembly ---------------
 * (js/spectral_scene.js:898-1003, 1188-1366). BLASes are built in local space by a
 * binned-SAH builder of this library (in place of three-mesh-bvh's MeshBVH, leaf
 * size <= 8) and flattened to the threaded layout; the TLAS is the reference's
 * median split with stable centroid sort (buildTlasRecords), leaf size 2. */
typedef struct sbgi_scene_builder sbgi_scene_builder;

int32_t sbgi_sb_create(sbgi_scene_builder **out);
/* Append one uber material record (28 floats); returns its index in *id. */
int32_t sbgi_sb_add_material(sbgi_scene_builder *b, const float *rec28, uint32_t *id);
/* n uber records at once (28 floats each); *first_id is the id of the first. */
int32_t sbgi_sb_add_materials(sbgi_scene_builder *b, const float *recs28, uint32_t n, uint32_t *first_id);
/* Add one local-space mesh and build its BLAS. normals/uvs may be NULL.
 * tri_material gives the uber index of every triangle (NULL = all 0). */
int32_t sbgi_sb_add_mesh(sbgi_scene_builder *b, const float *positions, uint32_t vcount,
                         const float *normals, const float *uvs, const uint32_t *indices,
                         uint32_t tcount, const uint32_t *tri_material, uint32_t *mesh_id);
/* Add one instance of a mesh; world is a column-major 4x4 (three.js Matrix4.elements). */
int32_t sbgi_sb_add_instance(sbgi_scene_builder *b, uint32_t mesh_id, const double *world16,
                             int32_t active, uint32_t *instance_id);
int32_t sbgi_sb_add_light(sbgi_scene_builder *b, const float *rec17);
/* Assemble pool, instance table and TLAS; *view points into builder-owned arrays
 * that stay valid until the builder is changed or destroyed. */
int32_t sbgi_sb_finalize(sbgi_scene_builder *b, sbgi_scene_view *view);
/* slot -> source instance id (the TLAS order), n = instance count */
int32_t sbgi_sb_tlas_order(const sbgi_scene_builder *b, uint32_t *order, uint32_t n);
/* Pool placement of a mesh after finalize (the blasList entry's vertBase / vertCount,
 * triangle base / count and blasRoot / blasEnd, js/spectral_scene.js:1140-1150):
 * out6 = {vertex_start, vertex_count, tri_start, tri_count, blas_root, blas_end}. */
int32_t sbgi_sb_mesh_ranges(const sbgi_scene_builder *b, uint32_t mesh_id, uint32_t *out6);
/* Replace the world matrix of a source instance and recompute the dynamic tail on the
 * host (updateTransforms): instance rows, det sign, world AABB, sparse TLAS refit
 * (refitFlatTlasDirty, js/gi_refit.js:352-430). Writes the TLAS slot of the instance. */
int32_t sbgi_sb_set_transform(sbgi_scene_builder *b, uint32_t instance_id,
                              const double *world16, uint32_t *slot);
const char *sbgi_sb_last_error(const sbgi_scene_builder *b);
int32_t sbgi_sb_destroy(sbgi_scene_builder *b);

#ifdef __cplusplus
}
#endif

#endif /* SBGI_H */
