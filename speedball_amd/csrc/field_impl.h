// field_impl.h — the probe-field object behind the C ABI, shared by the controller
// (field.hip) and the batched solve (batch.hip). Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "kernels.h"

namespace sbgi {

// Page-locked host storage for the mirrors that device edits copy back (BVH nodes,
// triangle rows): copies into it run at DMA rate instead of through a staging buffer.
template <class T>
struct PinnedAlloc {
    using value_type = T;
    PinnedAlloc() = default;
    template <class U>
    PinnedAlloc(const PinnedAlloc<U> &) {}
    T *allocate(size_t n) {
        void *p = nullptr;
        if (hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
        return static_cast<T *>(p);
    }
    void deallocate(T *p, size_t) { (void)hipHostFree(p); }
};
template <class T, class U>
bool operator==(const PinnedAlloc<T> &, const PinnedAlloc<U> &) { return true; }
template <class T, class U>
bool operator!=(const PinnedAlloc<T> &, const PinnedAlloc<U> &) { return false; }
template <class T>
using pinned_vector = std::vector<T, PinnedAlloc<T>>;

// message of the last failed sbgi_create on this thread
inline thread_local std::string g_create_error;

struct CascadeState {
    sbgi_cascade_desc d{};
    uint32_t total = 0, aw = 0, ah = 0, daw = 0, dah = 0, tile_d = 8;
    float max_dist = 0.f, cell_min = 0.f;
    float *irr = nullptr, *depth = nullptr, *meta = nullptr;
    ushort4 *atlas = nullptr, *datlas = nullptr;
    float *evis = nullptr;  // retained emitter visibility [probe * SBGI_MAX_EMITTERS + e]
    // reflection lobes (desc.reflections): rough state + atlas, glossy history + atlas
    float *rough = nullptr, *gnum = nullptr, *gden = nullptr;
    ushort4 *ratlas = nullptr, *gatlas = nullptr;
    uint32_t gtiles_x = 0, gtiles_y = 0, gaw = 0, gah = 0;
};

constexpr int kTicketRing = 16;

// A level-synchronous refit of a set of threaded BLAS ranges: the ranges' nodes ordered by
// height (leaves 0, interior = 1 + the taller child), level h = order[off[h], off[h + 1]).
struct RefitPlan {
    uint32_t *d_order = nullptr;
    size_t cap = 0;
    std::vector<uint32_t> off, key, h_order;
};

}  // namespace sbgi

struct sbgi_field {
    using CascadeState = sbgi::CascadeState;
    using RefitPlan = sbgi::RefitPlan;
    sbgi_field_desc desc{};
    hipStream_t stream = nullptr;
    CascadeState casc[SBGI_MAX_CASCADES];
    uint32_t max_total = 0;
    // scene (device)
    uint32_t *d_nodes = nullptr, *d_tri_index = nullptr, *d_tri_mat = nullptr;
    float *d_vdata = nullptr, *d_mats = nullptr, *d_lights = nullptr, *d_inst_bounds = nullptr;
    float4 *d_tri = nullptr, *d_inst = nullptr;
    uint4 *d_tlas = nullptr;
    int32_t *d_tlas_parent = nullptr;
    uint32_t *d_tlas_arrive = nullptr;
    // walk layouts: the BLAS pool in the walk's node format (1 copy, or 8 near-first
    // octant copies in MT_ORDERED mode) with the pool node behind each copy node, and
    // the SAH TLAS of MT_ORDERED mode
    uint32_t *d_walk_nodes = nullptr, *d_walk_perm = nullptr;
    uint4 *d_tlas_sah = nullptr;
    uint4 *d_tlas_oct = nullptr;  // its near-first copies per ray octant (closest-hit walk)
    uint32_t tlas_sah_count = 0;
    // per-frame TLAS on the device: world matrices (double, column-major 4x4 per slot),
    // the inflated bounds the SAH build reads, its scratch, and the staging area of
    // instance-record uploads {worlds f64[16 n], records f32[28 n], slots u32[n]}
    double *d_world = nullptr;
    float *d_inst_pad = nullptr;
    void *d_tlas_scratch = nullptr;
    void *d_inst_stage = nullptr;
    // probe classification scratch: 32 directions, then 32 hit records per probe
    float4 *d_classify = nullptr;
    size_t classify_cap = 0;
    // BLAS rebuild scratch (sort keys, depth queue, walk-copy positions), kept between
    // rebuilds and grown when a larger rebuild arrives
    void *d_rebuild = nullptr;
    size_t rebuild_bytes = 0;
    // skinned vertex sets (sbgi_set_skins): pool vertex per skinned vertex, rest pose
    // {pos, 0}, {normal, 0}, joints (4 x u16), weights, the bone palette, the leaves of the
    // skinned BLASes and their node / triangle ranges
    uint32_t *d_skin_vtx = nullptr;
    float4 *d_skin_rest = nullptr, *d_skin_w = nullptr, *d_bones = nullptr;
    uint2 *d_skin_joints = nullptr, *d_skin_nranges = nullptr, *d_skin_tranges = nullptr;
    uint32_t skin_vtx_n = 0, skin_bones_n = 0, skin_ranges_n = 0, skin_tranges_n = 0;
    uint32_t skin_nmax = 0, skin_tmax = 0;
    // bounds of d_nodes are newer than h_nodes after a device refit (structure is equal)
    bool nodes_dirty = false;
    // pinned upload staging ring (sbgi_set_instance_transforms, sbgi_skin)
    static constexpr int kStageSlots = 4;
    void *stage_host[kStageSlots] = {};
    size_t stage_cap[kStageSlots] = {};
    hipEvent_t stage_done[kStageSlots] = {};
    int stage_next = 0;
    // device timing of the last scene edits: kev[0..2] before skinning, after it, after the
    // BLAS refit and walk-copy update (sbgi_skin, sbgi_refit); kev[3..4] around the
    // instance-record scatter, instance bounds and TLAS (sbgi_set_instance_transforms, and
    // sbgi_skin / sbgi_refit when they update the TLAS)
    hipEvent_t kev[5] = {};
    bool kev_blas = false, kev_tlas = false;
    // level-synchronous refit plans: the skinned BLASes, and the ranges of the last sbgi_refit
    RefitPlan skin_plan, call_plan;
    // scene (host mirror, reference layouts)
    sbgi::pinned_vector<uint32_t> h_nodes, h_tri_index, h_tri_mat;
    std::vector<float> h_vdata, h_mats, h_lights;
    std::vector<double> h_world;  // per TLAS slot, column-major 4x4
    sbgi_scene_view view{};
    bool has_scene = false;
    // clustered NEE (SBGI_FIELD_CLUSTERED): cell lists over cascade 0's cells, rebuilt
    // whenever the light records change
    uint32_t *d_light_grid = nullptr;
    uint32_t lg_cells[3] = {0, 0, 0};
    // light-space candidate lists of the shadow walk (k_shadow.hip), rebuilt on the stream
    // whenever instance boxes or light records change; sl.list is null when the field walks
    // its shadow rays over the TLAS (SBGI_FIELD_NO_SHADOW_LISTS, or too large a build)
    ShadowLists sl{};
    // the lists predate the last change of boxes or lights: builds are skipped while the
    // walk tuning has them off and caught up when it turns them on
    bool sl_stale = false;
    // solve scratch
    float4 *rays = nullptr, *dirs = nullptr;
    int2 *hit_ids = nullptr;
    TraceScratch tw{};
    size_t queue_capacity = 0;
    void *tables = nullptr;
    BlendTables bt{};
    Counters *counters = nullptr;
    bool count_mode = false;
    uint32_t last_updated = 0, last_cascade = 0;
    // tickets and timing
    hipEvent_t ring[sbgi::kTicketRing] = {};
    hipEvent_t ev[5] = {};
    hipEvent_t tev[4] = {};  // trace sub-stages
    hipEvent_t eev = nullptr;  // end of the trace, before the emitter visibility rays
    // the captured tick (one hipGraph per frame) and the kernel arguments it was captured with
    hipGraph_t tick_graph = nullptr;
    hipGraphExec_t tick_exec = nullptr;
    std::vector<unsigned char> tick_key, scratch_key;
    uint64_t tick_replays = 0;
    bool tick_graph_failed = false, last_replayed = false;
    // stage events in captured ticks (sbgi_set_stage_timing), and whether the last tick's
    // stage events are current
    bool stage_timing = false, last_timed = false;
    // reflection lobe weight tables (desc.reflections != OFF)
    void *rtables = nullptr;
    ReflTables rt{};
    // receiver sampling (sbgi_sample_irradiance): staging of points, normals and results
    float *d_sample = nullptr;
    size_t sample_cap = 0;
    hipEvent_t sev[2] = {};
    sbgi_walk_tuning tuning{kDefaultTriRatio, kDefaultInstRatio, kDefaultRefillMin, kDefaultFetchChunk, 1u, 8u};
    // differentiable mode (sbgi_diff_*): parameters, the first update's FP32 atlas, the
    // gradients and the adjoint's scratch; the frozen hits and visibility masks are the
    // trace scratch tw of the last diff forward (diff_valid until any other solve)
    float *d_albedo = nullptr, *d_light_scale = nullptr, *d_gE = nullptr, *d_galb = nullptr, *d_glight = nullptr;
    int2 *d_albedo_map = nullptr;
    float4 *d_atlas1 = nullptr;
    void *d_adj_scratch = nullptr;
    uint32_t diff_albedo_n = 0;
    std::vector<int32_t> h_albedo_map;
    bool diff_ready = false, diff_valid = false;
    sbgi_tick_params diff_p{};
    hipEvent_t dev[4] = {};
    float ms_diff_fwd = 0.f, ms_diff_adj = 0.f;
    uint64_t ticket = 0;
    bool stats_valid = false;
    std::string err;
};

namespace sbgi {

inline int32_t fail(sbgi_field *f, int32_t code, const std::string &msg) {
    if (f)
        f->err = msg;
    else
        g_create_error = msg;
    return code;
}

inline int32_t hip_fail(sbgi_field *f, hipError_t e, const char *what) {
    return fail(f, SBGI_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_OK(f, expr)                                          \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) return hip_fail((f), e_, #expr);   \
    } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void **>(p), bytes);
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// (padding zeroed: batched solves compare these structs byte for byte)
inline DevCascade make_cascade(const sbgi_field *f, const CascadeState &cs, uint32_t offset, uint32_t updated) {
    DevCascade c;
    memset(&c, 0, sizeof(c));
    for (int a = 0; a < 3; a++) {
        c.gmin[a] = cs.d.grid_min[a];
        c.gsize[a] = cs.d.grid_size[a];
        c.res[a] = cs.d.res[a];
    }
    c.total = cs.total;
    c.offset = offset;
    c.updated = updated;
    c.rays = f->desc.rays_per_probe;
    c.depth_oct = f->desc.depth_oct_res;
    c.max_dist = cs.max_dist;
    c.cell_min = cs.cell_min;
    c.aw = cs.aw;
    c.ah = cs.ah;
    c.daw = cs.daw;
    c.dah = cs.dah;
    c.rough = f->desc.reflections != SBGI_REFLECT_OFF;
    return c;
}

// glossy tile (oct + 2 border) and update interval of the reflection quality
inline uint32_t glossy_tile(uint32_t q) { return q == SBGI_REFLECT_ULTRA ? 18u : (q == SBGI_REFLECT_HIGH ? 10u : 0u); }
inline uint32_t glossy_interval(uint32_t q) { return q == SBGI_REFLECT_HIGH ? 2u : 1u; }

inline ReflState refl_state(const CascadeState &cs) {
    return ReflState{cs.rough, cs.ratlas, cs.gnum, cs.gden, cs.gatlas, cs.gtiles_x, cs.gaw, cs.gah};
}

// MT_ORDERED and WATERTIGHT walk the SAH TLAS and the near-first octant copies
inline bool ordered(const sbgi_field *f) {
    return f->desc.traversal == SBGI_TRAV_MT_ORDERED || f->desc.traversal == SBGI_TRAV_WATERTIGHT;
}

inline DevScene make_scene(const sbgi_field *f) {
    DevScene s;
    memset(&s, 0, sizeof(s));
    s.blas = reinterpret_cast<const uint4 *>(f->d_nodes);
    s.nodes = reinterpret_cast<const uint4 *>(f->d_walk_nodes);
    s.blas_octants = ordered(f) ? 8u : 1u;
    s.shadow_octants = std::min(s.blas_octants, f->tuning.shadow_copies);
    s.tri = f->d_tri;
    s.tlas = ordered(f) ? f->d_tlas_sah : f->d_tlas;
    s.tlas_oct = ordered(f) ? f->d_tlas_oct : nullptr;
    s.inst = f->d_inst;
    s.mats = f->d_mats;
    s.tri_mat = f->d_tri_mat;
    s.lights = f->d_lights;
    s.node_count = f->view.node_count;
    s.tlas_count = ordered(f) ? f->tlas_sah_count : f->view.tlas_node_count;
    s.light_count = f->view.light_count;
    s.tri_count = f->view.tri_count;
    s.inst_count = f->view.inst_count;
    s.tri_ratio = f->tuning.tri_ratio;
    s.inst_ratio = f->tuning.inst_ratio;
    s.refill_min = f->tuning.refill_min;
    s.fetch_chunk = f->tuning.fetch_chunk;
    s.light_grid = f->d_light_grid;
    for (int a = 0; a < 3; a++) {
        s.lg_min[a] = f->casc[0].d.grid_min[a];
        s.lg_size[a] = f->casc[0].d.grid_size[a];
        s.lg_cells[a] = f->lg_cells[a];
    }
    s.emitter_base = f->view.emitter_base;
    s.emitter_count = f->view.emitter_count;
    s.watertight = f->desc.traversal == SBGI_TRAV_WATERTIGHT ? 1u : 0u;
    s.tri_index = f->d_tri_index;
    s.vdata = f->d_vdata;
    s.lds_inst = s.tlas_count * 2;
    if (f->sl.list && f->tuning.shadow_lists) {
        s.sl_proj = f->sl.proj;
        s.sl_cells = f->sl.cells;
        s.sl_list = f->sl.list;
        s.inst_box = f->sl.boxes;
    }
    return s;
}

inline size_t light_grid_cells(const sbgi_field *f) { return (size_t)f->lg_cells[0] * f->lg_cells[1] * f->lg_cells[2]; }

// the blend's emitter injection for one cascade solve
inline EmitterInject emitter_inject(const sbgi_field *f, const CascadeState &cs, const sbgi_tick_params &p) {
    return EmitterInject{f->d_lights, cs.meta, cs.evis, p.classify_strength, f->view.emitter_base,
                         f->has_scene ? f->view.emitter_count : 0u};
}

// the reflection atlases' clear (the rough and glossy clears of clearAtlasKernel /
// clearGlossyAtlasKernel, js/gi_probes.js:2690-2724): transparent black
inline hipError_t clear_refl_atlases(const CascadeState &cs, hipStream_t st) {
    hipError_t e = hipSuccess;
    if (cs.ratlas) e = launch_fill_atlas(cs.ratlas, (size_t)cs.aw * cs.ah, make_ushort4(0, 0, 0, 0), st);
    if (e == hipSuccess && cs.gatlas) e = launch_fill_atlas(cs.gatlas, (size_t)cs.gaw * cs.gah, make_ushort4(0, 0, 0, 0), st);
    return e;
}

// rough lobe (reads the depth moments before the blend rewrites them), the irradiance and
// depth blend, then the glossy cache
inline hipError_t launch_reflect_and_blend(sbgi_field *f, const CascadeState &cs, const DevCascade &c,
                                           const sbgi_tick_params &p, const float4 *rays, hipStream_t st) {
    const EmitterInject em = emitter_inject(f, cs, p);
    const ReflState rs = refl_state(cs);
    hipError_t e = hipSuccess;
    if (f->rtables) e = launch_rough(c, p, rays, f->rt, f->desc.blend_precision, cs.depth, rs, em, st);
    if (e == hipSuccess) e = launch_blend(c, p, rays, f->bt, f->desc.blend_precision, cs.irr, cs.depth, em, st);
    if (e == hipSuccess && f->rtables && f->rt.gtile)
        e = launch_glossy(c, p, rays, f->rt, f->desc.blend_precision, glossy_interval(f->desc.reflections), rs, em, st);
    return e;
}

}  // namespace sbgi
