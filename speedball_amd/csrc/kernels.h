// kernels.h — host launchers of the HIP kernels (one translation unit each).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "sbgi_dev.h"

// k_trace.hip
// Dense queue of shadow rays: rec[2i] = {origin.xyz, tmax}, rec[2i+1] = {dir.xyz,
// bits((ray << 6) | light)}. count points at the solve counters (kTraceCounterBytes):
// word 0 is the number of records appended this solve; the fetch counters of the primary
// (0), shadow (1), query (2) and emitter-visibility (3) walks start at word
// trace_heads(walk), one per 128 bytes.
constexpr size_t kTraceCounterBytes = 48 * 1024;
constexpr size_t trace_heads(uint32_t walk) { return (1u + (size_t)walk * kFetchHeads) * kHeadStride; }
static_assert((trace_heads(4)) * sizeof(uint32_t) <= kTraceCounterBytes, "solve counters overflow");
// Shadow-ray queue: 16-byte records {origin.xyz, (ray << 6) | light}; the walk re-derives
// direction and extent from the light record.
struct ShadowQueue {
    float4 *rec;
    uint32_t *count;
    uint32_t capacity;
};
// Per-solve trace scratch: closest-hit records {t, tri, inst, -} and per-ray light
// visibility masks (bit l set when the shadow ray toward light l is unblocked).
struct TraceScratch {
    float4 *hits;
    unsigned long long *vis;
    ShadowQueue queue;
};
constexpr size_t kTraceLdsBudget = 64 * 1024;  // TLAS + instances staged in LDS up to this size
size_t trace_lds_bytes(uint32_t tlas_count, uint32_t inst_count);
// probe classification (classifyKernel): CLASSIFY_RAYS fixed rays per probe and the
// buried fraction above which a probe is inactive (js/gi_probes.js:104-105)
constexpr uint32_t kClassifyRays = 32;
constexpr float kBackfaceFraction = 0.25f;
hipError_t launch_query(const DevScene &s, const float4 *in, float4 *out, uint32_t n, uint32_t *head, hipStream_t st);
hipError_t launch_raydirs(float4 *dirs, uint32_t rays, float jitter, hipStream_t st);
// How a solve marks a stage boundary with an event: recorded on the stream (fn null), or
// handed to fn, which the tick-graph builder uses to close the captured piece and place the
// event as a graph node between pieces.
struct StageMarker {
    hipError_t (*fn)(void *ctx, hipEvent_t e, hipStream_t st) = nullptr;
    void *ctx = nullptr;
    hipError_t operator()(hipEvent_t e, hipStream_t st) const { return fn ? fn(ctx, e, st) : hipEventRecord(e, st); }
};

// ev (optional): 4 events marked before the closest-hit walk, after it, after the
// shadow-ray emission and after the shadow walk.
// prev_atlas_f32 (optional): an FP32 bounce source used instead of the RGBA16F prev_atlas
// (the differentiable mode).
hipError_t launch_trace(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p, const float4 *dirs,
                        const float *meta, const ushort4 *prev_atlas, const float4 *prev_atlas_f32,
                        const TraceScratch &w, float4 *rays, int2 *hit_ids, Counters *counters, const hipEvent_t *ev,
                        hipStream_t st, const StageMarker &mark = StageMarker{});

// Batched trace (batch.hip): one segment per (field, cascade) solve of a frame. The two
// walks of a frame are single persistent launches over all segments; emission and
// shading run per segment.
struct TraceSeg {
    DevScene s;
    DevCascade c;
    float classify;
    const float4 *dirs;
    const float *meta;
    TraceScratch w;   // this segment's hits, visibility masks and shadow queue
    uint32_t *heads;  // this segment's fetch counters: primary walk, then shadow walk
    float bias;       // the cascade's hit bias (hit_bias), for the shadow walk's rays
};
// device walk tables of n segments (primary, then shadow) serialised into blob, which the
// caller uploads to a buffer of walk_tables_bytes(n) bytes
size_t walk_tables_bytes(uint32_t n_seg);
void walk_tables(const TraceSeg *segs, uint32_t n_seg, std::vector<unsigned char> &blob);
// mode 0: closest-hit walk over every segment's probe rays; 1: any-hit walk over every
// segment's shadow queue. prefix[j] = probe rays before segment j (n_seg + 1 entries, on
// the device) spreads the waves' first segments by work; done[n_seg] flags are cleared by
// the caller. block: 256 or 512 threads; waves: resident waves per CU (0: occupancy limit).
// shared: the scene of every segment when they all walk one (the cascades of one field):
// its TLAS and instance records are then staged in each block's LDS; else null.
hipError_t launch_walk_batch(int mode, const void *tables, uint32_t n_seg, const uint32_t *prefix, uint32_t total,
                             uint32_t *done, uint32_t block, uint32_t waves, bool wt, const DevScene *shared,
                             hipStream_t st);
hipError_t launch_emit(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p, const float4 *dirs,
                       const float *meta, const float4 *hits, const ShadowQueue &q, hipStream_t st);
hipError_t launch_resolve(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p, const float4 *dirs,
                          const float *meta, const ushort4 *prev_atlas, const float4 *hits,
                          const unsigned long long *vis, float4 *rays, hipStream_t st);
// emitterVisKernel: one any-hit ray per (updated probe, emitter of s); vis[probe * 16 + e]
// <- mix(sample, vis, 0.8). scratch: 2 float4 per ray (updated x emitters x 32 bytes, free
// between the shading and the next solve: the trace's hit records). heads: the walk's
// kFetchHeads fetch counters, zero on entry.
hipError_t launch_emitter_vis(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p, const float *meta,
                              float *vis, float4 *scratch, uint32_t *heads, hipStream_t st);

// classifyKernel over every probe of the cascade: kClassifyRays fixed closest-hit rays per
// probe from its grid position (dirs: kClassifyRays float4, written here; hits: total x
// kClassifyRays float4 of scratch), then meta[p] = {state, relocation offset.xyz}. heads:
// kFetchHeads fetch counters (cleared here).
hipError_t launch_classify(const DevScene &s, const DevCascade &c, float reloc_clamp, float4 *dirs, float4 *hits,
                           float *meta, uint32_t *heads, hipStream_t st);

// k_sample.hip: the receiver gather at n points (sbgi_sample_irradiance)
struct SampleCascadeIn {
    DevCascade c;
    const ushort4 *atlas, *datlas;
    const float *meta;
};
hipError_t launch_sample(const SampleCascadeIn *casc, uint32_t n_casc, const sbgi_sample_params &p, const float *pos,
                         const float *nvis, const float *ndir, uint32_t n, float *out, hipStream_t st);

// k_shadow.hip: light-space candidate lists of the shadow-ray walk (sbgi_dev.h
// shadow_cell). Device buffers: proj float4[lights * kShadowProj]; rects u32[lights *
// slots * 6], offsets u32[lights * kShadowCells] and bsum u32[shadow_list_blocks(lights)]
// (build scratch); total u32[1] (the
// entries written, identity list included); cells uint2[lights * kShadowRecs]; list
// u16[capacity >= slots + 1]: the identity list [0, slots), the NaN box's entry `slots`,
// then the cells' lists; boxes f32[(slots + 1) * 6], the boxes the list steps test.
struct ShadowLists {
    float4 *proj;
    float *boxes;
    uint32_t *rects, *offsets, *bsum, *total;
    uint2 *cells;
    uint16_t *list;
    uint16_t *order;  // per light, the slots nearest the light first ([lights][slots])
    uint32_t capacity;
};
// blocks of 256 cells that the build scans (the length of ShadowLists::bsum)
constexpr uint32_t shadow_list_blocks(uint32_t n_lights) { return (n_lights * kShadowCells + 255) / 256; }
// inst_box: the padded world box of every slot (6 floats), as the SAH TLAS build reads them
hipError_t launch_shadow_lists(const float *lights, uint32_t n_lights, const float *inst_box, uint32_t n_inst,
                               const ShadowLists &sl, hipStream_t st);

// k_lights.hip
// lightGridKernel: the cell lists {count, ids[16]} of s's lights over cells[0] x cells[1]
// x cells[2] cells spanning gmin + gsize (count 17 = overflow)
hipError_t launch_light_grid(const float *lights, uint32_t light_count, const float *gmin, const float *gsize,
                             const uint32_t *cells, uint32_t *grid, hipStream_t st);

// k_blend.hip
struct BlendTables {
    float *w_irr_t;     // [R][36]   cosine weights, k-major
    float *w_d_t;       // [R][TD]   depth weights, k-major
    float *norm_irr;    // [36]      sum_k cw
    float *norm_d;      // [TD]      sum_k dw
    uint16_t *frag;     // MFMA A fragments (bf16/fp16 hi and lo), see k_blend.hip
    float *norm_irr_q;  // [36] normalisers of the quantized weights
    float *norm_d_q;    // [TD]
};
size_t blend_tables_bytes(uint32_t rays, uint32_t depth_texels);
void blend_tables_carve(void *base, uint32_t rays, uint32_t depth_texels, BlendTables &t);
hipError_t launch_weights(const DevCascade &c, const sbgi_tick_params &p, const float4 *dirs, const BlendTables &t,
                          uint32_t precision, hipStream_t st);
// Emitters injected into the irradiance gather as virtual rays (injectEmitterVirtualRays):
// lights[(base + e) * 17] for e < count, their retained visibility vis[probe * 16 + e], and
// the probe trace origins (meta relocation x classify). count 0: no injection.
struct EmitterInject {
    const float *lights;
    const float *meta;
    const float *vis;
    float classify;
    uint32_t base, count;
};
hipError_t launch_blend(const DevCascade &c, const sbgi_tick_params &p, const float4 *rays, const BlendTables &t,
                        uint32_t precision, float *irr, float *depth, const EmitterInject &em, hipStream_t st);

// k_reflect.hip: the rough and glossy reflection lobes
struct ReflTables {
    float *w_rough_t;                   // [R][36] max(dir_t . dir_k, 0)^8, k-major
    float *norm_rough, *norm_rough_q;   // [36] sum_k W (exact, and of the quantized weights)
    uint16_t *frag_rough;               // MFMA A fragments, hi then lo
    float *w_gloss_t;                   // [R][G*G] ^64 over the glossy tile's texels
    float *norm_gloss, *norm_gloss_q;   // [G*G]
    uint16_t *frag_gloss;
    uint32_t gtile;                     // glossy tile G (oct + 2), 0 without the glossy cache
};
// one cascade's reflection state: rough {rgb x coverage, coverage} per irradiance texel
// and its atlas; glossy numerator / support per tile texel and the glossy atlas
struct ReflState {
    float *rough;
    ushort4 *ratlas;
    float *gnum, *gden;
    ushort4 *gatlas;
    uint32_t gtiles_x, gaw, gah;
};
size_t refl_tables_bytes(uint32_t rays, uint32_t gtile);
void refl_tables_carve(void *base, uint32_t rays, uint32_t gtile, ReflTables &t);
hipError_t launch_weights_refl(uint32_t rays, const float4 *dirs, const ReflTables &t, uint32_t precision,
                               hipStream_t st);
// before the irradiance blend: it reads the depth moments as the probe's seeded sentinel
hipError_t launch_rough(const DevCascade &c, const sbgi_tick_params &p, const float4 *rays, const ReflTables &t,
                        uint32_t precision, const float *depth, const ReflState &s, const EmitterInject &em,
                        hipStream_t st);
// interval: 1 solves every glossy texel; 2 the texels with index % 2 == p.glossy_phase
hipError_t launch_glossy(const DevCascade &c, const sbgi_tick_params &p, const float4 *rays, const ReflTables &t,
                         uint32_t precision, uint32_t interval, const ReflState &s, const EmitterInject &em,
                         hipStream_t st);
hipError_t launch_fill_atlas(ushort4 *atlas, size_t texels, ushort4 v, hipStream_t st);

// k_upload.hip
// atlas_f32 (optional): also store the irradiance tiles unrounded (differentiable mode)
hipError_t launch_upload(const DevCascade &c, const sbgi_tick_params &p, const float *irr, const float *depth,
                         ushort4 *atlas, ushort4 *datlas, float4 *atlas_f32, hipStream_t st,
                         const float *rough = nullptr, ushort4 *ratlas = nullptr);
hipError_t launch_clear_atlas(ushort4 *atlas, size_t texels, hipStream_t st);

// k_scene.hip
hipError_t launch_pack_tris(const uint32_t *tri_index, const float *vdata, const uint32_t *tri_mat, const float *mats,
                            uint32_t tri_count, float4 *tri, hipStream_t st);
hipError_t launch_pack_tlas(const float *mats, uint32_t tlas_base, uint32_t tlas_count, uint4 *tlas, hipStream_t st);
hipError_t launch_pack_inst(const float *mats, uint32_t inst_base, uint32_t inst_count, float4 *inst, hipStream_t st);
// walk-format node copies: bounds of copy node i <- pool node perm[i] (after a refit)
hipError_t launch_oct_gather(uint32_t *oct_nodes, const uint32_t *nodes, const uint32_t *perm, uint32_t n_total,
                             hipStream_t st);
hipError_t launch_refit_tlas(float *mats, uint32_t tlas_base, uint32_t tlas_count, const float *inst_bounds,
                             const int32_t *parent, uint32_t *arrive, hipStream_t st);
// ranged variants: triangles / walk-copy nodes of the listed [x, y) ranges (max_len = the
// longest range), and the refit climbing from a list of leaf nodes
hipError_t launch_pack_tris_ranges(const uint32_t *tri_index, const float *vdata, const uint32_t *tri_mat,
                                   const float *mats, uint32_t tri_count, const uint2 *ranges, uint32_t n_ranges,
                                   uint32_t max_len, float4 *tri, hipStream_t st);
hipError_t launch_oct_gather_ranges(uint32_t *oct_nodes, const uint32_t *nodes, const uint32_t *perm, uint32_t n_nodes,
                                    uint32_t copies, const uint2 *ranges, uint32_t n_ranges, uint32_t max_len,
                                    hipStream_t st);
// level-synchronous BLAS refit: order lists nodes by height, level h is
// order[level_off[h], level_off[h + 1]) (level_off on the host, levels + 1 entries)
hipError_t launch_refit_levels(uint32_t *nodes, const uint32_t *tri_index, const float *vdata, const uint32_t *order,
                               const uint32_t *level_off, uint32_t levels, hipStream_t st);
// the same over the rebuild's device depth table (deepest depth first)
hipError_t launch_refit_depths(uint32_t *nodes, const uint32_t *tri_index, const float *vdata, const uint32_t *order,
                               const uint32_t *lv, uint32_t depths, uint32_t cap, const uint32_t *blocks,
                               hipStream_t st);

// k_tlas.hip — instance world bounds (and the inflated copy the SAH build reads) and the
// one-workgroup SAH TLAS build: tlas receives 2n - 1 walk-format nodes
size_t tlas_build_scratch_bytes(uint32_t n);
hipError_t launch_inst_bounds(const float *mats, uint32_t inst_base, uint32_t n, const uint32_t *nodes,
                              const double *world, float *bounds, float *padded, hipStream_t st);
hipError_t launch_build_tlas(const float *padded, uint32_t n, void *scratch, uint4 *tlas, hipStream_t st);
// 8 near-first copies (one per ray direction octant) of a walk-format TLAS of m nodes,
// out: 8 m nodes; TLASes of more than kTlasOctMax nodes get none
constexpr uint32_t kTlasOctMax = 1023;
hipError_t launch_tlas_octants(const uint4 *tlas, uint32_t m, uint4 *out, hipStream_t st);

// k_build.hip — device BLAS rebuild (see that file). ranges: {root, N, t0, T} per BLAS;
// tri_prefix: triangles before each range (n_ranges + 1, on the device).
size_t rebuild_sort_bytes(uint32_t total);
hipError_t launch_rebuild_prepare(const uint4 *ranges, const uint32_t *tri_prefix, uint32_t n_ranges, uint32_t total,
                                  const uint32_t *tri_index, const float *vdata, float4 *cen, uint32_t *cbounds,
                                  unsigned long long *keys, uint32_t *vals, unsigned long long *keys_sorted,
                                  uint32_t *vals_sorted, void *sort_tmp, size_t sort_bytes, uint32_t end_bit,
                                  hipStream_t st);
hipError_t launch_permute_tris(const uint4 *ranges, const uint32_t *tri_prefix, uint32_t n_ranges, uint32_t total,
                               const uint32_t *sorted_vals, const uint32_t *old_index, const uint32_t *old_mat,
                               uint32_t *tri_index, uint32_t *tri_mat, hipStream_t st);
// top-down build by depth over a device depth table lv (see k_build.hip): tasks holds cap
// entries with the roots first; blocks[d] is the grid of depth d's launch
uint32_t rebuild_depth_bound(uint32_t max_leaves);
hipError_t launch_build_depths(const uint4 *ranges, const uint32_t *tri_prefix, const unsigned long long *keys,
                               uint4 *tasks, uint32_t cap, uint32_t *lv, uint32_t depths, const uint32_t *blocks,
                               uint32_t *nodes, uint32_t *order, hipStream_t st);
hipError_t launch_copy_roots(const uint4 *ranges, uint32_t n_ranges, uint32_t copies, uint32_t *qpos, uint32_t qstride,
                             hipStream_t st);
hipError_t launch_copy_positions(const uint32_t *nodes, const uint32_t *order, const uint32_t *lv, uint32_t depths,
                                 uint32_t cap, const uint32_t *blocks, uint32_t copies, uint32_t *qpos,
                                 uint32_t qstride, hipStream_t st);
hipError_t launch_write_copies(const uint32_t *nodes, const uint32_t *order, uint32_t n, const uint32_t *n_built,
                               uint32_t copies, const uint32_t *qpos, uint32_t qstride, uint32_t node_count,
                               uint32_t *walk, uint32_t *perm, hipStream_t st);

// k_dynamic.hip — device skinning, staged instance-record scatter, fill
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint32_t n, hipStream_t st);
hipError_t launch_skin(const uint32_t *vtx, const float4 *rest, const uint2 *joints, const float4 *weights,
                       const float4 *bones, uint32_t n, float *vdata, hipStream_t st);
hipError_t launch_scatter_inst(const uint32_t *slots, const float *recs, const double *worlds, uint32_t n, float *mats,
                               uint32_t inst_base, double *world, float4 *inst, hipStream_t st);

// k_adjoint.hip — adjoint of the differentiable two-update solve (see that file). w holds
// the forward's frozen hits and visibility masks, atlas1 the first update's FP32 atlas;
// gE2 and gE1 are P x 36 x 3; g_albedo (n x 3) and g_light (lights) accumulate.
size_t adjoint_scratch_bytes(uint32_t rays_total, uint32_t cells, uint32_t lights);
hipError_t launch_adjoint(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p, const float4 *dirs,
                          const BlendTables &t, const TraceScratch &w, const float4 *atlas1, const float *gE2,
                          float *gE1, float *g_albedo, float *g_light, void *scratch, hipStream_t st);
