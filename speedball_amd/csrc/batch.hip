// batch.hip — batched solves of many probe fields (sbgi_batch_* of include/sbgi.h).
//
// A batch groups fields of uneven size (their own scenes, cascades and ray counts) on one
// device. One frame updates every cascade of every field: each (field, cascade) solve is a
// segment with its own trace scratch, and the frame runs
//     per field: ray directions, blend weights (and atlas clears when asked)
//     one clear of every segment's visibility masks and fetch counters
//     ONE persistent closest-hit walk over all segments (device-side work queue)
//     per segment: shadow-ray emission
//     ONE persistent any-hit walk over all segments' shadow queues
//     per segment: shading, blend, upload
// on the batch's stream. The frame is captured into a hipGraph and replayed while every
// kernel argument stays the same (scene pointers, cascades, parameters, schedule); a frame
// whose arguments differ (the first one, a clear, new parameters or scene edits that moved
// a buffer) runs eagerly and is captured again. Fields in a batch are updated only through
// it; their own streams are joined before and after every frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "field_impl.h"
#include "kernels.h"

using namespace sbgi;

struct sbgi_batch {
    std::vector<sbgi_field *> fields;
    hipStream_t stream = nullptr;
    int device = 0;
    // segments: (field, cascade) with their scratch
    struct Seg {
        uint32_t field, cascade, rays_total, lights;
    };
    std::vector<Seg> segs;
    void *d_scratch = nullptr;   // hits | vis | queue records | segment rays, per segment
    uint32_t *d_ctr = nullptr;   // per segment: queue count + fetch heads (kTraceCounterBytes)
    uint32_t *d_done = nullptr;  // per segment: done flags of the two walks
    uint32_t *d_prefix = nullptr;
    void *d_tables = nullptr;
    std::vector<float4 *> hits, rays;
    std::vector<unsigned long long *> vis;
    std::vector<float4 *> qrec;
    std::vector<uint32_t> qcap;
    uint32_t total_rays = 0;
    uint32_t walk_block = 512;
    uint32_t walk_waves = 0;  // resident waves per CU of the batched walks (0: occupancy limit)
    // capture state: the argument blob of the captured frame
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> captured_key, tables_blob;
    hipEvent_t ev[2] = {};
    hipEvent_t join = nullptr;
    bool replayed = false, timed = false;
    uint64_t ticket = 0;
    std::string err;
};

namespace {

int32_t bfail(sbgi_batch *b, int32_t code, const std::string &msg) {
    if (b) b->err = msg;
    return code;
}

#define BHIP_OK(b, expr)                                                                       \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return bfail((b), SBGI_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void drop_graph(sbgi_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->captured_key.clear();
}

// The frame's segments as the kernels see them (full-field solves of every cascade).
void frame_segments(const sbgi_batch *b, const sbgi_tick_params *params, std::vector<TraceSeg> &out) {
    out.clear();
    for (size_t j = 0; j < b->segs.size(); j++) {
        const auto &g = b->segs[j];
        const sbgi_field *f = b->fields[g.field];
        const CascadeState &cs = f->casc[g.cascade];
        TraceSeg t;
        memset(&t, 0, sizeof(t));
        t.s = make_scene(f);
        t.c = make_cascade(f, cs, 0, cs.total);
        t.classify = params[g.field].classify_strength;
        t.bias = hit_bias(t.c.cell_min, params[g.field].dbg_trace_bias, params[g.field].dbg_ray_eps);
        t.dirs = f->dirs;
        t.meta = cs.meta;
        t.w.hits = b->hits[j];
        t.w.vis = b->vis[j];
        t.w.queue.rec = b->qrec[j];
        t.w.queue.capacity = b->qcap[j];
        t.w.queue.count = b->d_ctr + (size_t)j * (kTraceCounterBytes / 4);
        t.heads = t.w.queue.count + trace_heads(0);
        out.push_back(t);
    }
}

// Enqueue one frame on the batch stream (eagerly or under capture).
int32_t enqueue_frame(sbgi_batch *b, const sbgi_tick_params *params, const std::vector<TraceSeg> &ts) {
    hipStream_t st = b->stream;
    for (size_t i = 0; i < b->fields.size(); i++) {
        sbgi_field *f = b->fields[i];
        const sbgi_tick_params &p = params[i];
        const DevCascade c0 = make_cascade(f, f->casc[0], 0, f->casc[0].total);
        if (p.needs_clear)
            for (uint32_t c = 0; c < f->desc.cascade_count; c++) {
                const CascadeState &cs = f->casc[c];
                BHIP_OK(b, launch_clear_atlas(cs.atlas, (size_t)cs.aw * cs.ah, st));
                BHIP_OK(b, launch_clear_atlas(cs.datlas, (size_t)cs.daw * cs.dah, st));
                BHIP_OK(b, launch_fill_u32(reinterpret_cast<uint32_t *>(cs.evis), 0u, cs.total * SBGI_MAX_EMITTERS, st));
                BHIP_OK(b, clear_refl_atlases(cs, st));
            }
        BHIP_OK(b, launch_raydirs(f->dirs, c0.rays, p.frame_jitter, st));
        BHIP_OK(b, launch_weights(c0, p, f->dirs, f->bt, f->desc.blend_precision, st));
        if (f->rtables) BHIP_OK(b, launch_weights_refl(c0.rays, f->dirs, f->rt, f->desc.blend_precision, st));
    }
    const uint32_t n_seg = (uint32_t)b->segs.size();
    // visibility masks, queue counts, fetch heads and done flags of every segment
    BHIP_OK(b, launch_fill_u32(reinterpret_cast<uint32_t *>(b->vis[0]), 0u, 2u * b->total_rays, st));
    BHIP_OK(b, launch_fill_u32(b->d_ctr, 0u, n_seg * (uint32_t)(kTraceCounterBytes / 4), st));
    BHIP_OK(b, launch_fill_u32(b->d_done, 0u, 2u * n_seg, st));
    bool wt = false;
    for (const sbgi_field *f : b->fields) wt = wt || f->desc.traversal == SBGI_TRAV_WATERTIGHT;
    // one scene for every segment (the cascades of one field): the walks stage it in LDS
    const DevScene *shared = n_seg ? &ts[0].s : nullptr;
    for (uint32_t j = 1; j < n_seg && shared; j++)
        if (memcmp(&ts[j].s, &ts[0].s, sizeof(DevScene)) != 0) shared = nullptr;
    BHIP_OK(b, launch_walk_batch(0, b->d_tables, n_seg, b->d_prefix, b->total_rays, b->d_done, b->walk_block,
                                 b->walk_waves, wt, shared, st));
    for (uint32_t j = 0; j < n_seg; j++) {
        const sbgi_tick_params &p = params[b->segs[j].field];
        BHIP_OK(b, launch_emit(ts[j].s, ts[j].c, p, ts[j].dirs, ts[j].meta, ts[j].w.hits, ts[j].w.queue, st));
    }
    BHIP_OK(b, launch_walk_batch(1, b->d_tables, n_seg, b->d_prefix, b->total_rays, b->d_done + n_seg, b->walk_block,
                                 b->walk_waves, wt, shared, st));
    for (uint32_t j = 0; j < n_seg; j++) {
        const auto &g = b->segs[j];
        sbgi_field *f = b->fields[g.field];
        const CascadeState &cs = f->casc[g.cascade];
        const sbgi_tick_params &p = params[g.field];
        BHIP_OK(b, launch_resolve(ts[j].s, ts[j].c, p, ts[j].dirs, ts[j].meta, cs.atlas, ts[j].w.hits, ts[j].w.vis,
                                  b->rays[j], st));
        BHIP_OK(b, launch_emitter_vis(ts[j].s, ts[j].c, p, ts[j].meta, cs.evis, ts[j].w.hits,
                                      ts[j].w.queue.count + trace_heads(3), st));
        BHIP_OK(b, launch_reflect_and_blend(f, cs, ts[j].c, p, b->rays[j], st));
        BHIP_OK(b, launch_upload(ts[j].c, p, cs.irr, cs.depth, cs.atlas, cs.datlas, nullptr, st, cs.rough, cs.ratlas));
    }
    return SBGI_OK;
}

// Every by-value kernel argument of a frame: replaying a graph is valid only while this is
// unchanged.
void frame_key(const sbgi_batch *b, const sbgi_tick_params *params, const std::vector<TraceSeg> &ts,
               std::vector<unsigned char> &key) {
    key.clear();
    auto put = [&](const void *p, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        key.insert(key.end(), c, c + n);
    };
    put(params, sizeof(sbgi_tick_params) * b->fields.size());
    put(ts.data(), sizeof(TraceSeg) * ts.size());
    put(&b->walk_block, sizeof(b->walk_block));
    put(&b->walk_waves, sizeof(b->walk_waves));
    for (const sbgi_field *f : b->fields) put(&f->bt, sizeof(f->bt));
}

int32_t run_frame(sbgi_batch *b, const sbgi_tick_params *params, bool allow_graph) {
    std::vector<TraceSeg> ts;
    frame_segments(b, params, ts);
    std::vector<unsigned char> key;
    frame_key(b, params, ts, key);
    bool clear = false;
    for (size_t i = 0; i < b->fields.size(); i++) clear = clear || params[i].needs_clear;
    // walk tables: rewritten only when the frame's arguments changed (a replay reads them)
    if (key != b->captured_key) {
        drop_graph(b);
        walk_tables(ts.data(), (uint32_t)ts.size(), b->tables_blob);
        BHIP_OK(b, hipMemcpyAsync(b->d_tables, b->tables_blob.data(), b->tables_blob.size(), hipMemcpyHostToDevice,
                                  b->stream));
    }
    BHIP_OK(b, hipEventRecord(b->ev[0], b->stream));
    int32_t rc = SBGI_OK;
    if (b->exec && key == b->captured_key) {
        BHIP_OK(b, hipGraphLaunch(b->exec, b->stream));
        b->replayed = true;
    } else if (allow_graph && !clear) {
        BHIP_OK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_frame(b, params, ts);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc != SBGI_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        BHIP_OK(b, e);
        b->graph = g;
        BHIP_OK(b, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        BHIP_OK(b, hipGraphLaunch(b->exec, b->stream));
        b->captured_key = key;
        b->replayed = false;
    } else {
        if ((rc = enqueue_frame(b, params, ts)) != SBGI_OK) return rc;
        b->replayed = false;
    }
    BHIP_OK(b, hipEventRecord(b->ev[1], b->stream));
    b->timed = true;
    return SBGI_OK;
}

// The segment scratch was sized at sbgi_batch_create: a field whose scene now has more
// lights (shadow-queue capacity) or whose solve size changed needs a new batch.
int32_t check_sizes(sbgi_batch *b) {
    for (size_t j = 0; j < b->segs.size(); j++) {
        const auto &g = b->segs[j];
        const sbgi_field *f = b->fields[g.field];
        if (!f->has_scene) return bfail(b, SBGI_E_STATE, "a field of the batch has no scene");
        if (std::max(1u, f->view.light_count) > g.lights ||
            (uint64_t)f->casc[g.cascade].total * f->desc.rays_per_probe != g.rays_total)
            return bfail(b, SBGI_E_STATE, "a field's scene or size changed since sbgi_batch_create: create the batch again");
    }
    return SBGI_OK;
}

}  // namespace

extern "C" {

int32_t sbgi_batch_create(sbgi_field *const *fields, uint32_t n, sbgi_batch **out) {
    if (!fields || !n || !out) return SBGI_E_INVALID;
    *out = nullptr;
    sbgi_batch *b = new (std::nothrow) sbgi_batch();
    if (!b) return SBGI_E_NOMEM;
    b->device = fields[0] ? fields[0]->desc.device : 0;
    for (uint32_t i = 0; i < n; i++) {
        sbgi_field *f = fields[i];
        if (!f || !f->has_scene || f->desc.device != b->device) {
            delete b;
            return SBGI_E_INVALID;
        }
        for (uint32_t k = 0; k < i; k++)
            if (fields[k] == f) {
                delete b;
                return SBGI_E_INVALID;
            }
        b->fields.push_back(f);
    }
    // segments and their scratch sizes
    size_t bytes = 0;
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; i++) {
        const sbgi_field *f = fields[i];
        for (uint32_t c = 0; c < f->desc.cascade_count; c++) {
            const uint64_t r = (uint64_t)f->casc[c].total * f->desc.rays_per_probe;
            const uint32_t L = std::max(1u, f->view.light_count);
            if (r * L >= (1ull << 32)) {
                delete b;
                return SBGI_E_CAPACITY;
            }
            b->segs.push_back({i, c, (uint32_t)r, L});
            total += r;
            bytes += r * (16 + 16) + r * L * 16;
        }
    }
    if (total >= (1ull << 31)) {
        delete b;
        return SBGI_E_CAPACITY;
    }
    b->total_rays = (uint32_t)total;
    const uint32_t n_seg = (uint32_t)b->segs.size();
    bytes += total * 8;  // visibility masks, one array for every segment
    auto cleanup = [&](int32_t code) {
        (void)sbgi_batch_destroy(b);
        return code;
    };
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess)
        return cleanup(SBGI_E_DEVICE);
    if (hipMalloc(&b->d_scratch, bytes) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&b->d_ctr), (size_t)n_seg * kTraceCounterBytes) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&b->d_done), (size_t)n_seg * 8) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&b->d_prefix), ((size_t)n_seg + 1) * 4) != hipSuccess ||
        hipMalloc(&b->d_tables, walk_tables_bytes(n_seg)) != hipSuccess)
        return cleanup(SBGI_E_NOMEM);
    // carve: the visibility masks first (one contiguous array), then per segment
    char *p = static_cast<char *>(b->d_scratch);
    unsigned long long *vis = reinterpret_cast<unsigned long long *>(p);
    p += total * 8;
    std::vector<uint32_t> prefix(n_seg + 1, 0u);
    for (uint32_t j = 0; j < n_seg; j++) {
        const auto &g = b->segs[j];
        b->vis.push_back(vis);
        vis += g.rays_total;
        b->hits.push_back(reinterpret_cast<float4 *>(p));
        p += (size_t)g.rays_total * 16;
        b->rays.push_back(reinterpret_cast<float4 *>(p));
        p += (size_t)g.rays_total * 16;
        b->qrec.push_back(reinterpret_cast<float4 *>(p));
        b->qcap.push_back(g.rays_total * g.lights);
        p += (size_t)g.rays_total * g.lights * 16;
        prefix[j + 1] = prefix[j] + g.rays_total;
    }
    if (hipMemcpy(b->d_prefix, prefix.data(), prefix.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(SBGI_E_DEVICE);
    *out = b;
    return SBGI_OK;
}

int32_t sbgi_batch_tick(sbgi_batch *b, const sbgi_tick_params *params, uint64_t *ticket) {
    if (!b || !params) return SBGI_E_INVALID;
    BHIP_OK(b, hipSetDevice(b->device));
    int32_t rc0 = check_sizes(b);
    if (rc0 != SBGI_OK) return rc0;
    for (sbgi_field *f : b->fields) {
        // scene edits enqueued on the field's stream come first
        BHIP_OK(b, hipEventRecord(b->join, f->stream));
        BHIP_OK(b, hipStreamWaitEvent(b->stream, b->join, 0));
        f->diff_valid = false;
        f->stats_valid = false;
    }
    int32_t rc = run_frame(b, params, true);
    if (rc != SBGI_OK) return rc;
    // and the field's later work comes after the frame
    BHIP_OK(b, hipEventRecord(b->join, b->stream));
    for (sbgi_field *f : b->fields) BHIP_OK(b, hipStreamWaitEvent(f->stream, b->join, 0));
    b->ticket++;
    if (ticket) *ticket = b->ticket;
    return SBGI_OK;
}

int32_t sbgi_batch_wait(sbgi_batch *b) {
    if (!b) return SBGI_E_INVALID;
    BHIP_OK(b, hipSetDevice(b->device));
    BHIP_OK(b, hipStreamSynchronize(b->stream));
    return SBGI_OK;
}

int32_t sbgi_batch_timing(sbgi_batch *b, float *ms, int32_t *replayed) {
    if (!b) return SBGI_E_INVALID;
    if (!b->timed) return bfail(b, SBGI_E_STATE, "no frame yet");
    BHIP_OK(b, hipEventSynchronize(b->ev[1]));
    float m = 0.f;
    BHIP_OK(b, hipEventElapsedTime(&m, b->ev[0], b->ev[1]));
    if (ms) *ms = m;
    if (replayed) *replayed = b->replayed ? 1 : 0;
    return SBGI_OK;
}

int32_t sbgi_batch_set_walk_block(sbgi_batch *b, uint32_t block) {
    if (!b || (block != 256 && block != 512)) return SBGI_E_INVALID;
    b->walk_block = block;
    return SBGI_OK;
}

int32_t sbgi_batch_set_walk_waves(sbgi_batch *b, uint32_t waves_per_cu) {
    if (!b || (waves_per_cu != 0 && (waves_per_cu < 4 || waves_per_cu > 64))) return SBGI_E_INVALID;
    b->walk_waves = waves_per_cu;
    return SBGI_OK;
}

int32_t sbgi_batch_get_walk_waves(const sbgi_batch *b, uint32_t *waves_per_cu) {
    if (!b || !waves_per_cu) return SBGI_E_INVALID;
    *waves_per_cu = b->walk_waves;
    return SBGI_OK;
}

int32_t sbgi_batch_autotune(sbgi_batch *b, const sbgi_tick_params *params, sbgi_walk_tuning *best,
                            uint32_t *best_block, float *best_ms) {
    if (!b || !params) return SBGI_E_INVALID;
    BHIP_OK(b, hipSetDevice(b->device));
    int32_t rc0 = check_sizes(b);
    if (rc0 != SBGI_OK) return rc0;
    for (sbgi_field *f : b->fields) {
        BHIP_OK(b, hipEventRecord(b->join, f->stream));
        BHIP_OK(b, hipStreamWaitEvent(b->stream, b->join, 0));
    }
    // one schedule for every field of the batch, the walk block ("tile") size and the
    // resident waves per CU; each candidate is timed on eager frames (the temporal state
    // advances, as with ticks)
    std::vector<sbgi_tick_params> p(params, params + b->fields.size());
    for (auto &q : p) q.needs_clear = 0;
    auto apply = [&](const sbgi_walk_tuning &t, uint32_t block, uint32_t waves) {
        for (sbgi_field *f : b->fields) {  // each field keeps its own choice of shadow walk
            const uint32_t lists = f->tuning.shadow_lists, copies = f->tuning.shadow_copies;
            f->tuning = t;
            f->tuning.shadow_lists = lists;
            f->tuning.shadow_copies = copies;
        }
        b->walk_block = block;
        b->walk_waves = waves;
    };
    auto time_it = [&](const sbgi_walk_tuning &t, uint32_t block, uint32_t waves, float &ms) -> int32_t {
        apply(t, block, waves);
        float best_run = 1e30f;
        for (int rep = 0; rep < 2; rep++) {
            int32_t rc = run_frame(b, p.data(), false);
            if (rc != SBGI_OK) return rc;
            BHIP_OK(b, hipEventSynchronize(b->ev[1]));
            float m = 0.f;
            BHIP_OK(b, hipEventElapsedTime(&m, b->ev[0], b->ev[1]));
            best_run = std::min(best_run, m);
        }
        ms = best_run;
        return SBGI_OK;
    };
    sbgi_walk_tuning cur = b->fields[0]->tuning;
    uint32_t block = b->walk_block, waves = b->walk_waves;
    float cur_ms = 0.f;
    int32_t rc = time_it(cur, block, waves, cur_ms);
    if (rc != SBGI_OK) return rc;
    static const uint32_t ratios[] = {2, 4, 8}, refills[] = {4, 8, 16}, chunks[] = {32, 64, 128}, blocks[] = {256, 512},
                          wave_caps[] = {0, 16, 8};
    for (int param = 0; param < 6; param++) {
        const int nc = param == 4 ? 2 : 3;
        for (int i = 0; i < nc; i++) {
            sbgi_walk_tuning t = cur;
            uint32_t bl = block, wv = waves;
            uint32_t v = 0;
            switch (param) {
                case 0: v = t.tri_ratio = ratios[i]; if (cur.tri_ratio == v) continue; break;
                case 1: v = t.inst_ratio = ratios[i]; if (cur.inst_ratio == v) continue; break;
                case 2: v = t.refill_min = refills[i]; if (cur.refill_min == v) continue; break;
                case 3: v = t.fetch_chunk = chunks[i]; if (cur.fetch_chunk == v) continue; break;
                case 4: bl = blocks[i]; if (bl == block) continue; break;
                default: wv = wave_caps[i]; if (wv == waves) continue; break;
            }
            float ms = 0.f;
            if ((rc = time_it(t, bl, wv, ms)) != SBGI_OK) return rc;
            if (ms < cur_ms * 0.99f) {
                cur = t;
                block = bl;
                waves = wv;
                cur_ms = ms;
            }
        }
    }
    apply(cur, block, waves);
    BHIP_OK(b, hipEventRecord(b->join, b->stream));
    for (sbgi_field *f : b->fields) BHIP_OK(b, hipStreamWaitEvent(f->stream, b->join, 0));
    cur.shadow_lists = 0;  // batch frames walk shadow rays over the TLAS
    cur.shadow_copies = b->fields[0]->tuning.shadow_copies;  // the fields keep their own
    if (best) *best = cur;
    if (best_block) *best_block = block;
    if (best_ms) *best_ms = cur_ms;
    return SBGI_OK;
}

const char *sbgi_batch_last_error(const sbgi_batch *b) { return b ? b->err.c_str() : "null batch"; }

int32_t sbgi_batch_destroy(sbgi_batch *b) {
    if (!b) return SBGI_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (b->d_scratch) (void)hipFree(b->d_scratch);
    if (b->d_ctr) (void)hipFree(b->d_ctr);
    if (b->d_done) (void)hipFree(b->d_done);
    if (b->d_prefix) (void)hipFree(b->d_prefix);
    if (b->d_tables) (void)hipFree(b->d_tables);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SBGI_OK;
}

}  // extern "C"
