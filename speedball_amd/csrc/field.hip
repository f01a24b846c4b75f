// field.hip — the probe-field controller behind the C ABI of include/sbgi.h.
//
// One sbgi_field owns one HIP stream on one device, the per-cascade temporal state
// (irradiance/depth moments, probe meta), the RGBA16F atlases, the ray scratch and
// the scene in device memory. A solve (sbgi_tick) enqueues
//     raydirs -> weights -> trace -> blend -> upload
// on the field's stream, the ordered pass that js/gi_probes.js:3836 submits through
// renderer.computeAsync, and returns a ticket; HIP events bracket every stage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "field_impl.h"
#include "kernels.h"
#include "sbgi_host.h"

using namespace sbgi;

// ---- device scratch of sbgi_query_rays_noalloc ---------------------------------------
// The walk's fetch counters, then the staged query records (32 B per ray) and results
// (16 B per ray) of at most SBGI_NOALLOC_MAX_RAYS rays: 776 KiB per GPU, held until the
// process ends. Host code reaches it only through the address that hipGetSymbolAddress
// returns on each device (g_scratch_dev): the array's own name is a host-side stand-in.
constexpr size_t kNoAllocHeadBytes = (size_t)kFetchHeads * kHeadStride * sizeof(uint32_t);
constexpr size_t kNoAllocBytes = kNoAllocHeadBytes + (size_t)SBGI_NOALLOC_MAX_RAYS * (32 + 16);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[kScratchDevices] = {};

// Once per device, with that device current: the device address of its scratch copy.
hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_dev[device] = static_cast<unsigned char *>(p);
    return hipSuccess;
}

unsigned char *scratch_of(int device) {
    if (device < 0 || device >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_dev[device];
}
}  // namespace

namespace {

void free_plan(RefitPlan &p) {
    dfree(p.d_order);
    p.cap = 0;
    p.off.clear();
    p.key.clear();
    p.h_order.clear();
}

// Build (or reuse, when the ranges are the same as last time) the height-ordered node list
// of n validated [root, end) ranges.
int32_t build_plan(sbgi_field *f, const uint32_t *ranges, uint32_t n, RefitPlan &p) {
    std::vector<uint32_t> key(ranges, ranges + 2 * (size_t)n);
    if (key == p.key && p.d_order) return SBGI_OK;
    std::vector<uint32_t> count;
    size_t total = 0;
    std::vector<std::vector<uint32_t>> heights(n);
    for (uint32_t r = 0; r < n; r++) {
        const uint32_t root = ranges[2 * r], end = ranges[2 * r + 1];
        std::vector<uint32_t> &H = heights[r];
        H.assign(end - root, 0u);
        for (uint32_t i = end; i-- > root;) {
            const uint32_t *nd = f->h_nodes.data() + (size_t)i * 8;
            uint32_t hh = 0;
            if (nd[7] == 0xFFFFFFFFu) {
                const uint32_t l = i + 1, rr = f->h_nodes[(size_t)l * 8 + 6];
                hh = std::max(H[l - root], H[rr - root]) + 1;
            }
            H[i - root] = hh;
            if (hh >= count.size()) count.resize(hh + 1, 0u);
            count[hh]++;
        }
        total += end - root;
    }
    p.off.assign(count.size() + 1, 0u);
    for (size_t k = 0; k < count.size(); k++) p.off[k + 1] = p.off[k] + count[k];
    p.h_order.assign(total, 0u);
    std::vector<uint32_t> fill(p.off.begin(), p.off.end() - 1);
    for (uint32_t r = 0; r < n; r++) {
        const uint32_t root = ranges[2 * r];
        for (size_t q = 0; q < heights[r].size(); q++) p.h_order[fill[heights[r][q]]++] = root + (uint32_t)q;
    }
    if (p.cap < total) {
        dfree(p.d_order);
        p.cap = 0;
        if (dalloc(&p.d_order, total * 4) != hipSuccess) return fail(f, SBGI_E_NOMEM, "refit plan allocation");
        p.cap = total;
    }
    HIP_OK(f, hipMemcpyAsync(p.d_order, p.h_order.data(), total * 4, hipMemcpyHostToDevice, f->stream));
    p.key = key;
    return SBGI_OK;
}

void free_skins(sbgi_field *f) {
    dfree(f->d_skin_vtx);
    dfree(f->d_skin_rest);
    dfree(f->d_skin_w);
    dfree(f->d_bones);
    dfree(f->d_skin_joints);
    dfree(f->d_skin_nranges);
    dfree(f->d_skin_tranges);
    f->skin_vtx_n = f->skin_bones_n = f->skin_ranges_n = f->skin_tranges_n = 0;
    free_plan(f->skin_plan);
    f->skin_nmax = f->skin_tmax = 0;
}

void free_shadow_lists(sbgi_field *f) {
    dfree(f->sl.proj);
    dfree(f->sl.boxes);
    dfree(f->sl.rects);
    dfree(f->sl.offsets);
    dfree(f->sl.bsum);
    dfree(f->sl.total);
    dfree(f->sl.cells);
    dfree(f->sl.list);
    dfree(f->sl.order);
    f->sl.capacity = 0;
}

void free_scene(sbgi_field *f) {
    free_shadow_lists(f);
    dfree(f->d_nodes);
    dfree(f->d_tri_index);
    dfree(f->d_tri_mat);
    dfree(f->d_vdata);
    dfree(f->d_mats);
    dfree(f->d_lights);
    dfree(f->d_inst_bounds);
    dfree(f->d_tri);
    dfree(f->d_inst);
    dfree(f->d_tlas);
    dfree(f->d_tlas_parent);
    free_plan(f->call_plan);
    dfree(f->d_tlas_arrive);
    dfree(f->tw.queue.rec);
    dfree(f->d_walk_nodes);
    dfree(f->d_walk_perm);
    dfree(f->d_tlas_sah);
    dfree(f->d_tlas_oct);
    dfree(f->d_world);
    dfree(f->d_inst_pad);
    dfree(f->d_tlas_scratch);
    dfree(f->d_inst_stage);
    dfree(f->d_rebuild);
    f->rebuild_bytes = 0;
    dfree(f->d_classify);
    f->classify_cap = 0;
    free_skins(f);
    dfree(f->d_albedo);
    dfree(f->d_light_scale);
    dfree(f->d_gE);
    dfree(f->d_galb);
    dfree(f->d_glight);
    dfree(f->d_albedo_map);
    dfree(f->d_atlas1);
    dfree(f->d_adj_scratch);
    f->diff_ready = f->diff_valid = false;
    f->tlas_sah_count = 0;
    f->tw.queue.capacity = 0;
    f->nodes_dirty = false;
    f->kev_blas = f->kev_tlas = false;
    f->has_scene = false;
}

bool is_int(float v) { return std::isfinite(v) && v == std::floor(v); }

// Validate the reference layouts so that every device access stays in bounds.
int32_t validate_scene(sbgi_field *f) {
    const sbgi_scene_view &v = f->view;
    if (v.node_count >= (1u << 24) || v.tri_count >= (1u << 24) || v.inst_count >= (1u << 24))
        return fail(f, SBGI_E_CAPACITY, "BVH pool or instance count exceeds the 24-bit index ceiling");
    if (v.light_count > SBGI_MAX_LIGHTS) return fail(f, SBGI_E_CAPACITY, "more than MAX_LIGHTS (64) lights");
    if (v.emitter_count > SBGI_MAX_EMITTERS) return fail(f, SBGI_E_CAPACITY, "more than 16 emitters");
    if ((uint64_t)v.emitter_base + v.emitter_count > v.light_count)
        return fail(f, SBGI_E_INVALID, "emitter records lie outside the light records");
    if (v.node_count == 0 || v.tri_count == 0 || v.inst_count == 0 || v.tlas_node_count == 0)
        return fail(f, SBGI_E_INVALID, "empty scene");
    if ((size_t)v.inst_base + (size_t)v.inst_count * 28 > v.tlas_base ||
        (size_t)v.tlas_base + (size_t)v.tlas_node_count * 12 > v.materials_len || v.inst_base % 28 != 0)
        return fail(f, SBGI_E_INVALID, "materials buffer too small for the instance/TLAS tail");
    const uint32_t n_mat = v.inst_base / 28;
    for (uint32_t i = 0; i < v.node_count; i++) {
        const uint32_t *nd = f->h_nodes.data() + (size_t)i * 8;
        if (!(nd[6] > i && nd[6] <= v.node_count)) return fail(f, SBGI_E_LAYOUT, "BLAS miss link out of order");
        if (nd[7] != 0xFFFFFFFFu) {
            uint64_t off = nd[7] & 0x00FFFFFFu, cnt = nd[7] >> 24;
            if (off + cnt > v.tri_count) return fail(f, SBGI_E_LAYOUT, "BLAS leaf triangle range out of bounds");
        }
    }
    for (uint32_t t = 0; t < v.tri_count; t++) {
        for (int k = 0; k < 3; k++)
            if (f->h_tri_index[(size_t)t * 3 + k] >= v.vertex_count)
                return fail(f, SBGI_E_LAYOUT, "triangle vertex index out of range");
        if (f->h_tri_mat[t] >= n_mat) return fail(f, SBGI_E_LAYOUT, "triangle material index out of range");
    }
    for (uint32_t i = 0; i < v.inst_count; i++) {
        const float *r = f->h_mats.data() + v.inst_base + (size_t)i * 28;
        if (!is_int(r[12]) || !is_int(r[13]) || !(r[12] >= 0.f && r[12] < r[13] && r[13] <= (float)v.node_count))
            return fail(f, SBGI_E_LAYOUT, "instance BLAS range out of bounds");
    }
    for (uint32_t i = 0; i < v.tlas_node_count; i++) {
        const float *r = f->h_mats.data() + v.tlas_base + (size_t)i * 12;
        if (!is_int(r[6]) || !(r[6] > (float)i && r[6] <= (float)v.tlas_node_count))
            return fail(f, SBGI_E_LAYOUT, "TLAS miss link out of order");
        if (r[8] >= 0.5f) {
            if (!is_int(r[7]) || !is_int(r[8]) || r[8] > 255.f || r[7] < 0.f ||
                r[7] + r[8] > (float)v.inst_count)
                return fail(f, SBGI_E_LAYOUT, "TLAS leaf instance range out of bounds");
        }
    }
    return SBGI_OK;
}

void build_tlas_parents(const float *tl, uint32_t n, std::vector<int32_t> &parent) {
    parent.assign(n, -1);
    for (uint32_t i = 0; i < n; i++) {
        if (tl[(size_t)i * 12 + 8] >= 0.5f) continue;
        uint32_t l = i + 1;
        if (l >= n) continue;
        uint32_t r = (uint32_t)tl[(size_t)l * 12 + 6];
        parent[l] = (int32_t)i;
        if (r < n) parent[r] = (int32_t)i;
    }
}

void invert_affine_rows(const float *rows12, double *world16) {
    // world = inverse of [rows | 0 0 0 1], column-major output
    double m[16] = {rows12[0], rows12[4], rows12[8], 0, rows12[1], rows12[5], rows12[9], 0,
                    rows12[2], rows12[6], rows12[10], 0, rows12[3], rows12[7], rows12[11], 1};
    double a = m[0], b = m[4], c = m[8], d = m[1], e = m[5], f = m[9], g = m[2], h = m[6], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    for (int k = 0; k < 16; k++) world16[k] = 0;
    world16[15] = 1;
    if (det == 0) return;
    double id = 1 / det;
    double inv[9] = {A * id, -(b * i - c * h) * id, (b * f - c * e) * id,
                     B * id, (a * i - c * g) * id,  -(a * f - c * d) * id,
                     C * id, -(a * h - b * g) * id, (a * e - b * d) * id};
    // inv is row-major 3x3
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) world16[q * 4 + r] = inv[r * 3 + q];
    for (int r = 0; r < 3; r++)
        world16[12 + r] = -(inv[r * 3 + 0] * m[12] + inv[r * 3 + 1] * m[13] + inv[r * 3 + 2] * m[14]);
}

// The shadow walk's candidate lists from the current instance boxes and light records
// (stream-ordered; a build that outgrows the list capacity leaves identity lists, which the
// next sized build fixes).
int32_t build_shadow_lists(sbgi_field *f) {
    if (!f->sl.list) return SBGI_OK;
    if (!f->tuning.shadow_lists) {  // not walked: built when the tuning turns them on
        f->sl_stale = true;
        return SBGI_OK;
    }
    HIP_OK(f, launch_shadow_lists(f->d_lights, f->view.light_count, f->d_inst_pad, f->view.inst_count, f->sl, f->stream));
    f->sl_stale = false;
    return SBGI_OK;
}

// Allocate the candidate lists of a new scene and build them with a capacity of twice the
// entries the scene needs now (one host sync), so that later edits rarely outgrow it. A
// scene whose build would cost more than kShadowListWork rectangle tests, or whose slots
// do not fit 16 bits, walks its shadow rays over the TLAS.
// Ticks take the lists by default (sbgi_walk_tuning::shadow_lists): nearest the light
// first, they cut the top-level steps of a shadow ray from 37 to 5 and the walk time by a
// fifth on the cache-resident instanced workload (2.5 MiB of node copies and triangles),
// and the frame by 3.5 % on the dynamic workload (320 MiB, BLAS reads from HBM; their
// rebuild every frame included), where in slot order they had cost 6.5 %.
constexpr uint64_t kShadowListWork = 1ull << 28;
constexpr size_t kShadowCopyScene = 16ull << 20;
int32_t size_shadow_lists(sbgi_field *f) {
    free_shadow_lists(f);
    const sbgi_scene_view &v = f->view;
    const size_t walk_bytes = (size_t)v.node_count * 32 * (ordered(f) ? 8 : 1) + (size_t)v.tri_count * 48;
    f->tuning.shadow_lists = 1u;
    // past it, shadow rays also walk one BLAS copy instead of their octant's near-first copy:
    // a shadow ray needs any occluder, and on the dynamic workload (8 copies: 269 MiB) its
    // walks take 8.49 instead of 8.82 ms a frame; on the cache-resident instanced workload
    // one copy costs the shadow walk 6 %
    f->tuning.shadow_copies = walk_bytes <= kShadowCopyScene ? 8u : 1u;
    const uint64_t work = (uint64_t)v.light_count * kShadowCells * v.inst_count;
    if ((f->desc.flags & SBGI_FIELD_NO_SHADOW_LISTS) || v.light_count == 0 || v.inst_count == 0 ||
        v.inst_count > 65535u || work > kShadowListWork)
        return SBGI_OK;
    const size_t nl = v.light_count, ni = v.inst_count;
    auto alloc_list = [&](size_t entries) {
        entries = (entries + 3) & ~(size_t)3;  // the walk reads entries four per 64-bit word
        dfree(f->sl.list);
        f->sl.capacity = (uint32_t)entries;
        return dalloc(&f->sl.list, entries * sizeof(uint16_t));
    };
    if (dalloc(&f->sl.proj, nl * kShadowProj * sizeof(float4)) != hipSuccess ||
        dalloc(&f->sl.boxes, (ni + 1) * 6 * sizeof(float)) != hipSuccess ||
        dalloc(&f->sl.rects, nl * ni * 6 * sizeof(uint32_t)) != hipSuccess ||
        dalloc(&f->sl.offsets, nl * kShadowCells * sizeof(uint32_t)) != hipSuccess ||
        dalloc(&f->sl.bsum, shadow_list_blocks((uint32_t)nl) * sizeof(uint32_t)) != hipSuccess ||
        dalloc(&f->sl.total, sizeof(uint32_t)) != hipSuccess ||
        dalloc(&f->sl.cells, nl * kShadowRecs * sizeof(uint2)) != hipSuccess ||
        dalloc(&f->sl.order, nl * ni * sizeof(uint16_t)) != hipSuccess || alloc_list(ni + (1u << 20)) != hipSuccess) {
        free_shadow_lists(f);
        return fail(f, SBGI_E_NOMEM, "shadow list allocation");
    }
    HIP_OK(f, launch_shadow_lists(f->d_lights, v.light_count, f->d_inst_pad, v.inst_count, f->sl, f->stream));
    uint32_t total = 0;
    HIP_OK(f, hipMemcpyAsync(&total, f->sl.total, sizeof(total), hipMemcpyDeviceToHost, f->stream));
    HIP_OK(f, hipStreamSynchronize(f->stream));
    const uint64_t want = ni + 1 + 2 * ((uint64_t)total - ni - 1) + 4096;
    if (want > f->sl.capacity) {
        if (want >= (1ull << 32) || alloc_list((size_t)want) != hipSuccess) {
            free_shadow_lists(f);
            return fail(f, SBGI_E_NOMEM, "shadow list allocation");
        }
        HIP_OK(f, launch_shadow_lists(f->d_lights, v.light_count, f->d_inst_pad, v.inst_count, f->sl, f->stream));
    }
    f->sl_stale = false;
    return SBGI_OK;
}

// Instance world bounds from the device node pool and world matrices, then the TLAS: in
// MT_ORDERED mode the SAH TLAS is rebuilt; in MT_COMPAT mode the reference TLAS is refit
// bottom-up when refit_compat is set (updateTransforms' refit, js/spectral_scene.js:
// 1402-1514; refitFlatTlasRange, js/gi_refit.js:245-292). Stream-ordered, no host sync.
int32_t device_tlas(sbgi_field *f, bool refit_compat) {
    const sbgi_scene_view &v = f->view;
    HIP_OK(f, launch_inst_bounds(f->d_mats, v.inst_base, v.inst_count, f->d_nodes, f->d_world, f->d_inst_bounds,
                                 f->d_inst_pad, f->stream));
    if (ordered(f)) {
        HIP_OK(f, launch_build_tlas(f->d_inst_pad, v.inst_count, f->d_tlas_scratch, f->d_tlas_sah, f->stream));
        if (f->d_tlas_oct) HIP_OK(f, launch_tlas_octants(f->d_tlas_sah, f->tlas_sah_count, f->d_tlas_oct, f->stream));
    } else if (refit_compat) {
        HIP_OK(f, launch_fill_u32(f->d_tlas_arrive, 0u, v.tlas_node_count, f->stream));
        HIP_OK(f, launch_refit_tlas(f->d_mats, v.tlas_base, v.tlas_node_count, f->d_inst_bounds, f->d_tlas_parent,
                                    f->d_tlas_arrive, f->stream));
        HIP_OK(f, launch_pack_tlas(f->d_mats, v.tlas_base, v.tlas_node_count, f->d_tlas, f->stream));
    }
    return build_shadow_lists(f);
}

// Bring the host mirror of the node pool up to date after device refits (before host code
// reads its bounds).
int32_t sync_nodes(sbgi_field *f) {
    if (!f->nodes_dirty) return SBGI_OK;
    HIP_OK(f, hipMemcpyAsync(f->h_nodes.data(), f->d_nodes, f->h_nodes.size() * 4, hipMemcpyDeviceToHost, f->stream));
    HIP_OK(f, hipStreamSynchronize(f->stream));
    f->nodes_dirty = false;
    return SBGI_OK;
}

// A pinned host buffer of at least `bytes` from the staging ring whose previous upload has
// completed; stage_commit records that completion after the upload is enqueued.
int32_t stage_acquire(sbgi_field *f, size_t bytes, void **out, int *slot) {
    const int s = f->stage_next;
    f->stage_next = (s + 1) % sbgi_field::kStageSlots;
    if (!f->stage_done[s]) HIP_OK(f, hipEventCreateWithFlags(&f->stage_done[s], hipEventDisableTiming));
    HIP_OK(f, hipEventSynchronize(f->stage_done[s]));
    if (f->stage_cap[s] < bytes) {
        if (f->stage_host[s]) HIP_OK(f, hipHostFree(f->stage_host[s]));
        f->stage_host[s] = nullptr;
        f->stage_cap[s] = 0;
        const size_t cap = std::max<size_t>(bytes, 64 * 1024);
        if (hipHostMalloc(&f->stage_host[s], cap, hipHostMallocDefault) != hipSuccess)
            return fail(f, SBGI_E_NOMEM, "pinned staging allocation");
        f->stage_cap[s] = cap;
    }
    *out = f->stage_host[s];
    *slot = s;
    return SBGI_OK;
}

int32_t stage_commit(sbgi_field *f, int slot) {
    HIP_OK(f, hipEventRecord(f->stage_done[slot], f->stream));
    return SBGI_OK;
}

int32_t record_kev(sbgi_field *f, int i) {
    if (!f->kev[i]) HIP_OK(f, hipEventCreate(&f->kev[i]));
    HIP_OK(f, hipEventRecord(f->kev[i], f->stream));
    return SBGI_OK;
}

int32_t upload_walk_nodes(sbgi_field *f) {
    const sbgi_scene_view &v = f->view;
    std::vector<std::pair<uint32_t, uint32_t>> roots;
    for (uint32_t i = 0; i < v.inst_count; i++) {
        const float *r = f->h_mats.data() + v.inst_base + (size_t)i * 28;
        roots.emplace_back((uint32_t)r[12], (uint32_t)r[13]);
    }
    std::sort(roots.begin(), roots.end());
    roots.erase(std::unique(roots.begin(), roots.end()), roots.end());
    for (size_t i = 1; ordered(f) && i < roots.size(); i++)
        if (roots[i].first < roots[i - 1].second) return fail(f, SBGI_E_LAYOUT, "overlapping instance BLAS ranges");
    std::vector<uint32_t> oc, perm;
    if (!sbgi::build_node_copies(f->h_nodes.data(), v.node_count, roots, ordered(f) ? 8u : 1u, oc, perm))
        return fail(f, SBGI_E_LAYOUT, "BLAS range is not a threaded subtree");
    dfree(f->d_walk_nodes);
    dfree(f->d_walk_perm);
    if (dalloc(&f->d_walk_nodes, oc.size() * 4) != hipSuccess || dalloc(&f->d_walk_perm, perm.size() * 4) != hipSuccess)
        return fail(f, SBGI_E_NOMEM, "walk node allocation");
    HIP_OK(f, hipMemcpyAsync(f->d_walk_nodes, oc.data(), oc.size() * 4, hipMemcpyHostToDevice, f->stream));
    HIP_OK(f, hipMemcpyAsync(f->d_walk_perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, f->stream));
    HIP_OK(f, hipStreamSynchronize(f->stream));
    return SBGI_OK;
}

int32_t repack_all(sbgi_field *f) {
    const sbgi_scene_view &v = f->view;
    HIP_OK(f, launch_pack_tris(f->d_tri_index, f->d_vdata, f->d_tri_mat, f->d_mats, v.tri_count, f->d_tri, f->stream));
    HIP_OK(f, launch_pack_tlas(f->d_mats, v.tlas_base, v.tlas_node_count, f->d_tlas, f->stream));
    HIP_OK(f, launch_pack_inst(f->d_mats, v.inst_base, v.inst_count, f->d_inst, f->stream));
    return SBGI_OK;
}

}  // namespace

extern "C" {

// clustered lighting: the cell lists of the current light records, stream-ordered
int32_t build_light_grid(sbgi_field *f) {
    if (!f->d_light_grid) return SBGI_OK;
    const sbgi_cascade_desc &c0 = f->casc[0].d;
    HIP_OK(f, launch_light_grid(f->d_lights, f->view.light_count, c0.grid_min, c0.grid_size, f->lg_cells,
                                f->d_light_grid, f->stream));
    return SBGI_OK;
}

void sbgi_tick_params_default(sbgi_tick_params *p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->frame_jitter = 0.f;
    p->hysteresis = 0.60f;  // JITTER_HYSTERESIS_DEFAULTS.gated (js/gi_probes.js:146-149)
    p->hysteresis_exponent = 1.f;
    p->depth_sharpness = 7.f;
    p->radiance_clamp = 8.f;
    p->classify_strength = 0.f;
    p->filter_strength = 1.f;
    p->filter_smooth = 0.5f;
    p->temp_change_sigma1 = 2.5f;
    p->temp_change_h_drop = 0.30f;
    p->temp_clamp_sigma = 6.f;
    p->nir_gate = 0.f;
    p->nir_gain = 1.f;
    p->sky_intensity = 1.f;
    p->dbg_trace_bias = 1.f;
    p->dbg_ray_eps = 1.f;
    p->dbg_direct = 1.f;
    p->dbg_emissive = 1.f;
    p->dbg_albedo = 1.f;
    p->dbg_bounce = 1.f;
    p->dbg_cosine_power = 1.f;
    p->dbg_noise_h_boost = 0.25f;
    p->dbg_change_sigma0 = 0.75f;
    p->dbg_min_change_h = 0.55f;
    p->dbg_var_eps = 0.000025f;
    p->dbg_var_rel = 0.0025f;
    p->dbg_depth_history = 1.f;
    p->dbg_filter_k = 1.f;
    p->dbg_filter_rel = 1.f;
    p->dbg_filter_eps = 1.f;
}

const char *sbgi_version(void) { return "sbgi 0.1.0 gfx950"; }

int32_t sbgi_create(const sbgi_field_desc *desc, sbgi_field **out) {
    if (!desc || !out) return fail(nullptr, SBGI_E_INVALID, "null argument");
    *out = nullptr;
    const uint32_t R = desc->rays_per_probe;
    if (R < SBGI_RAYS_MIN || R > SBGI_RAYS_MAX || R % 16 != 0)
        return fail(nullptr, SBGI_E_INVALID, "rays_per_probe must be a multiple of 16 in [32, 256]");
    if (desc->cascade_count < 1 || desc->cascade_count > SBGI_MAX_CASCADES)
        return fail(nullptr, SBGI_E_INVALID, "cascade_count must be 1..4");
    if (desc->depth_oct_res != 6 && desc->depth_oct_res != 14)
        return fail(nullptr, SBGI_E_INVALID, "depth_oct_res must be 6 or 14");
    if (desc->blend_precision > SBGI_BLEND_FP16) return fail(nullptr, SBGI_E_INVALID, "unknown blend precision");
    if (desc->traversal > SBGI_TRAV_WATERTIGHT)
        return fail(nullptr, SBGI_E_INVALID, "traversal must be MT_COMPAT, MT_ORDERED or WATERTIGHT");
    if (desc->reflections > SBGI_REFLECT_ULTRA) return fail(nullptr, SBGI_E_INVALID, "unknown reflection quality");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || desc->device < 0 || desc->device >= ndev)
        return fail(nullptr, SBGI_E_DEVICE, "HIP device not available");
    sbgi_field *f = new (std::nothrow) sbgi_field();
    if (!f) return fail(nullptr, SBGI_E_NOMEM, "out of host memory");
    f->desc = *desc;
    auto bail = [&](int32_t code) {
        g_create_error = f->err;
        sbgi_destroy(f);
        return code;
    };
    if (hipSetDevice(desc->device) != hipSuccess) return bail(fail(f, SBGI_E_DEVICE, "hipSetDevice failed"));
    if (scratch_setup(desc->device) != hipSuccess)
        return bail(fail(f, SBGI_E_DEVICE, "device address of the query scratch not available"));
    if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(f, SBGI_E_DEVICE, "hipStreamCreate failed"));
    for (auto &e : f->ring)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(f, SBGI_E_DEVICE, "hipEventCreate failed"));
    for (auto &e : f->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(f, SBGI_E_DEVICE, "hipEventCreate failed"));
    for (auto &e : f->tev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(f, SBGI_E_DEVICE, "hipEventCreate failed"));
    for (auto &e : f->dev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(f, SBGI_E_DEVICE, "hipEventCreate failed"));
    if (hipEventCreate(&f->eev) != hipSuccess) return bail(fail(f, SBGI_E_DEVICE, "hipEventCreate failed"));
    for (auto &e : f->sev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(f, SBGI_E_DEVICE, "hipEventCreate failed"));
    const uint32_t tile_d = desc->depth_oct_res + 2;
    for (uint32_t ci = 0; ci < desc->cascade_count; ci++) {
        CascadeState &cs = f->casc[ci];
        cs.d = desc->cascades[ci];
        const uint64_t total = (uint64_t)cs.d.res[0] * cs.d.res[1] * cs.d.res[2];
        if (cs.d.res[0] < 1 || cs.d.res[1] < 1 || cs.d.res[2] < 1 || total >= (1u << 24))
            return bail(fail(f, SBGI_E_INVALID, "cascade resolution out of range"));
        cs.total = (uint32_t)total;
        cs.tile_d = tile_d;
        cs.aw = cs.d.res[0] * 8;
        cs.ah = cs.d.res[1] * cs.d.res[2] * 8;
        cs.daw = desc->depth_oct_res == 6 ? cs.aw : cs.d.res[0] * tile_d;
        cs.dah = desc->depth_oct_res == 6 ? cs.ah : cs.d.res[1] * cs.d.res[2] * tile_d;
        // maxDist = |gridSize|, cellMin = max(1e-4, min cell) (js/gi_probes.js:3209-3211, 3361-3362)
        double gx = cs.d.grid_size[0], gy = cs.d.grid_size[1], gz = cs.d.grid_size[2];
        cs.max_dist = (float)std::sqrt(gx * gx + gy * gy + gz * gz);
        double cm = std::min({gx / std::max(1u, cs.d.res[0] - 1), gy / std::max(1u, cs.d.res[1] - 1),
                              gz / std::max(1u, cs.d.res[2] - 1)});
        cs.cell_min = (float)std::max(1e-4, cm);
        if (dalloc(&cs.evis, (size_t)cs.total * SBGI_MAX_EMITTERS * 4) != hipSuccess)
            return bail(fail(f, SBGI_E_NOMEM, "device allocation failed"));
        (void)hipMemsetAsync(cs.evis, 0, (size_t)cs.total * SBGI_MAX_EMITTERS * 4, f->stream);
        if (desc->reflections != SBGI_REFLECT_OFF) {
            const size_t G = glossy_tile(desc->reflections), gt = (size_t)cs.total * G * G;
            cs.gtiles_x = G ? std::max(1u, (uint32_t)std::ceil(std::sqrt((double)cs.total))) : 0;
            cs.gtiles_y = G ? (cs.total + cs.gtiles_x - 1) / cs.gtiles_x : 0;
            cs.gaw = cs.gtiles_x * (uint32_t)G;
            cs.gah = cs.gtiles_y * (uint32_t)G;
            const size_t ra = (size_t)cs.d.res[0] * 8 * cs.d.res[1] * cs.d.res[2] * 8;
            if (dalloc(&cs.rough, (size_t)cs.total * 64 * 16) != hipSuccess || dalloc(&cs.ratlas, ra * 8) != hipSuccess ||
                (G && (dalloc(&cs.gnum, gt * 16) != hipSuccess || dalloc(&cs.gden, gt * 4) != hipSuccess ||
                       dalloc(&cs.gatlas, (size_t)cs.gaw * cs.gah * 8) != hipSuccess)))
                return bail(fail(f, SBGI_E_NOMEM, "device allocation failed"));
            (void)hipMemsetAsync(cs.rough, 0, (size_t)cs.total * 64 * 16, f->stream);
            (void)hipMemsetAsync(cs.ratlas, 0, ra * 8, f->stream);
            if (G) {
                (void)hipMemsetAsync(cs.gnum, 0, gt * 16, f->stream);
                (void)hipMemsetAsync(cs.gden, 0, gt * 4, f->stream);
                (void)hipMemsetAsync(cs.gatlas, 0, (size_t)cs.gaw * cs.gah * 8, f->stream);
            }
        }
        const size_t depth_texels = desc->depth_oct_res == 6 ? 64 : (size_t)tile_d * tile_d;
        if (dalloc(&cs.irr, (size_t)cs.total * 64 * 16) != hipSuccess ||
            dalloc(&cs.depth, (size_t)cs.total * depth_texels * 8) != hipSuccess ||
            dalloc(&cs.meta, (size_t)cs.total * 16) != hipSuccess ||
            dalloc(&cs.atlas, (size_t)cs.aw * cs.ah * 8) != hipSuccess ||
            dalloc(&cs.datlas, (size_t)cs.daw * cs.dah * 8) != hipSuccess)
            return bail(fail(f, SBGI_E_NOMEM, "device allocation failed"));
        (void)hipMemsetAsync(cs.irr, 0, (size_t)cs.total * 64 * 16, f->stream);
        (void)hipMemsetAsync(cs.depth, 0, (size_t)cs.total * depth_texels * 8, f->stream);
        (void)hipMemsetAsync(cs.meta, 0, (size_t)cs.total * 16, f->stream);
        if (launch_clear_atlas(cs.atlas, (size_t)cs.aw * cs.ah, f->stream) != hipSuccess ||
            launch_clear_atlas(cs.datlas, (size_t)cs.daw * cs.dah, f->stream) != hipSuccess)
            return bail(fail(f, SBGI_E_DEVICE, "atlas clear failed"));
        f->max_total = std::max(f->max_total, cs.total);
    }
    if (desc->flags & SBGI_FIELD_CLUSTERED) {
        const uint32_t *r0 = desc->cascades[0].res;
        if (r0[0] < 2 || r0[1] < 2 || r0[2] < 2)
            return bail(fail(f, SBGI_E_INVALID, "clustered lighting needs cascade 0 res >= 2 on every axis"));
        for (int a = 0; a < 3; a++) f->lg_cells[a] = r0[a] - 1;
        if (dalloc(&f->d_light_grid, light_grid_cells(f) * (SBGI_LIGHTS_PER_CELL + 1) * 4) != hipSuccess)
            return bail(fail(f, SBGI_E_NOMEM, "device allocation failed"));
        (void)hipMemsetAsync(f->d_light_grid, 0, light_grid_cells(f) * (SBGI_LIGHTS_PER_CELL + 1) * 4, f->stream);
    }
    const uint32_t tdn = desc->depth_oct_res == 6 ? 36 : desc->depth_oct_res * desc->depth_oct_res;
    const size_t tb = blend_tables_bytes(R, tdn);
    if (dalloc(&f->rays, (size_t)f->max_total * R * 16) != hipSuccess || dalloc(&f->dirs, 256 * 16) != hipSuccess ||
        dalloc(&f->tables, tb) != hipSuccess || dalloc(&f->counters, sizeof(Counters)) != hipSuccess)
        return bail(fail(f, SBGI_E_NOMEM, "device allocation failed"));
    if (dalloc(&f->tw.hits, (size_t)f->max_total * R * 16) != hipSuccess ||
        dalloc(&f->tw.vis, (size_t)f->max_total * R * 8) != hipSuccess ||
        dalloc(&f->tw.queue.count, kTraceCounterBytes) != hipSuccess)
        return bail(fail(f, SBGI_E_NOMEM, "device allocation failed"));
    if ((uint64_t)f->max_total * R >= (1ull << 26))
        return bail(fail(f, SBGI_E_CAPACITY, "probes x rays per solve must stay below 2^26"));
    if (desc->flags & SBGI_FIELD_KEEP_HIT_IDS) {
        if (dalloc(&f->hit_ids, (size_t)f->max_total * R * 8) != hipSuccess)
            return bail(fail(f, SBGI_E_NOMEM, "device allocation failed"));
    }
    if (hipMemsetAsync(f->tables, 0, tb, f->stream) != hipSuccess ||
        hipMemsetAsync(f->counters, 0, sizeof(Counters), f->stream) != hipSuccess)
        return bail(fail(f, SBGI_E_DEVICE, "clearing the weight tables failed"));
    blend_tables_carve(f->tables, R, tdn, f->bt);
    if (desc->reflections != SBGI_REFLECT_OFF) {
        const size_t rb = refl_tables_bytes(R, glossy_tile(desc->reflections));
        if (dalloc(&f->rtables, rb) != hipSuccess) return bail(fail(f, SBGI_E_NOMEM, "device allocation failed"));
        if (hipMemsetAsync(f->rtables, 0, rb, f->stream) != hipSuccess)
            return bail(fail(f, SBGI_E_DEVICE, "clearing the reflection tables failed"));
        refl_tables_carve(f->rtables, R, glossy_tile(desc->reflections), f->rt);
    }
    if (hipStreamSynchronize(f->stream) != hipSuccess) return bail(fail(f, SBGI_E_DEVICE, "stream sync failed"));
    *out = f;
    return SBGI_OK;
}

int32_t sbgi_set_scene(sbgi_field *f, const sbgi_scene_view *s) {
    if (!f || !s) return SBGI_E_INVALID;
    if (!s->bvh_nodes || !s->tri_index || !s->vertex_data || !s->tri_material || !s->materials ||
        (s->light_count && !s->lights))
        return fail(f, SBGI_E_INVALID, "scene view has a null buffer");
    HIP_OK(f, hipSetDevice(f->desc.device));
    HIP_OK(f, hipStreamSynchronize(f->stream));
    f->view = *s;
    f->h_nodes.assign(s->bvh_nodes, s->bvh_nodes + (size_t)s->node_count * 8);
    f->h_tri_index.assign(s->tri_index, s->tri_index + (size_t)s->tri_count * 3);
    f->h_tri_mat.assign(s->tri_material, s->tri_material + s->tri_count);
    f->h_vdata.assign(s->vertex_data, s->vertex_data + (size_t)s->vertex_count * 8);
    f->h_mats.assign(s->materials, s->materials + s->materials_len);
    f->h_lights.assign(s->lights ? s->lights : nullptr, s->lights ? s->lights + (size_t)s->light_count * 17 : nullptr);
    int32_t rc = validate_scene(f);
    if (rc != SBGI_OK) {
        f->has_scene = false;
        return rc;
    }
    free_scene(f);
    const sbgi_scene_view &v = f->view;
    std::vector<int32_t> tparent;
    build_tlas_parents(f->h_mats.data() + v.tlas_base, v.tlas_node_count, tparent);
    // world matrices recovered from the inverse rows; exact ones arrive via sbgi_set_instance_transforms
    f->h_world.assign((size_t)v.inst_count * 16, 0.0);
    for (uint32_t i = 0; i < v.inst_count; i++)
        invert_affine_rows(f->h_mats.data() + v.inst_base + (size_t)i * 28, f->h_world.data() + (size_t)i * 16);
    const size_t stage_bytes = (size_t)v.inst_count * (16 * sizeof(double) + 28 * sizeof(float) + sizeof(uint32_t));
    if (dalloc(&f->d_nodes, f->h_nodes.size() * 4) != hipSuccess ||
        dalloc(&f->d_tri_index, f->h_tri_index.size() * 4) != hipSuccess ||
        dalloc(&f->d_tri_mat, f->h_tri_mat.size() * 4) != hipSuccess ||
        dalloc(&f->d_vdata, f->h_vdata.size() * 4) != hipSuccess || dalloc(&f->d_mats, f->h_mats.size() * 4) != hipSuccess ||
        dalloc(&f->d_lights, std::max<size_t>(17, f->h_lights.size()) * 4) != hipSuccess ||
        dalloc(&f->d_inst_bounds, (size_t)v.inst_count * 24) != hipSuccess ||
        dalloc(&f->d_inst_pad, (size_t)v.inst_count * 24) != hipSuccess ||
        dalloc(&f->d_world, f->h_world.size() * sizeof(double)) != hipSuccess ||
        dalloc(&f->d_inst_stage, stage_bytes) != hipSuccess ||
        (ordered(f) && (dalloc(&f->d_tlas_sah, (2 * (size_t)v.inst_count - 1) * 32) != hipSuccess ||
                        dalloc(&f->d_tlas_scratch, tlas_build_scratch_bytes(v.inst_count)) != hipSuccess)) ||
        (ordered(f) && 2 * (size_t)v.inst_count - 1 <= kTlasOctMax &&
         dalloc(&f->d_tlas_oct, 8 * (2 * (size_t)v.inst_count - 1) * 32) != hipSuccess) ||
        dalloc(&f->d_tri, (size_t)v.tri_count * 48) != hipSuccess ||
        dalloc(&f->d_inst, (size_t)v.inst_count * 64) != hipSuccess ||
        dalloc(&f->d_tlas, (size_t)v.tlas_node_count * 32) != hipSuccess ||
        dalloc(&f->d_tlas_parent, (size_t)v.tlas_node_count * 4) != hipSuccess ||
        dalloc(&f->d_tlas_arrive, (size_t)v.tlas_node_count * 4) != hipSuccess) {
        free_scene(f);
        return fail(f, SBGI_E_NOMEM, "device allocation for the scene failed");
    }
    // shadow queue: at most one record per (ray, light)
    const size_t cap = (size_t)f->max_total * f->desc.rays_per_probe * std::max(1u, v.light_count);
    if (cap >= (1ull << 32) || dalloc(&f->tw.queue.rec, cap * 16) != hipSuccess) {
        free_scene(f);
        return fail(f, SBGI_E_NOMEM, "device allocation for the shadow-ray queue failed");
    }
    f->tw.queue.capacity = (uint32_t)cap;
    auto up = [&](void *d, const void *h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, f->stream) : hipSuccess;
    };
    HIP_OK(f, up(f->d_nodes, f->h_nodes.data(), f->h_nodes.size() * 4));
    HIP_OK(f, up(f->d_tri_index, f->h_tri_index.data(), f->h_tri_index.size() * 4));
    HIP_OK(f, up(f->d_tri_mat, f->h_tri_mat.data(), f->h_tri_mat.size() * 4));
    HIP_OK(f, up(f->d_vdata, f->h_vdata.data(), f->h_vdata.size() * 4));
    HIP_OK(f, up(f->d_mats, f->h_mats.data(), f->h_mats.size() * 4));
    HIP_OK(f, up(f->d_lights, f->h_lights.data(), f->h_lights.size() * 4));
    if ((rc = build_light_grid(f)) != SBGI_OK) return rc;
    HIP_OK(f, up(f->d_world, f->h_world.data(), f->h_world.size() * sizeof(double)));
    HIP_OK(f, up(f->d_tlas_parent, tparent.data(), tparent.size() * 4));
    rc = repack_all(f);
    if (rc != SBGI_OK) return rc;
    // instance bounds for later refits; in MT_ORDERED mode also the first SAH TLAS. The
    // reference TLAS of MT_COMPAT mode is used as given.
    f->tlas_sah_count = ordered(f) ? 2 * v.inst_count - 1 : 0;
    if ((rc = device_tlas(f, false)) != SBGI_OK) return rc;
    if ((rc = size_shadow_lists(f)) != SBGI_OK) return rc;
    HIP_OK(f, hipStreamSynchronize(f->stream));
    if ((rc = upload_walk_nodes(f)) != SBGI_OK) {
        free_scene(f);
        return rc;
    }
    f->has_scene = true;
    return SBGI_OK;
}

int32_t sbgi_update_ranges(sbgi_field *f, int32_t buf_id, const uint32_t *ranges, uint32_t n_ranges, const void *src) {
    if (!f || (n_ranges && (!ranges || !src))) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    void *dev = nullptr;
    void *host = nullptr;
    size_t elems = 0, esz = 4;
    switch (buf_id) {
        case SBGI_BUF_BVH_NODES: dev = f->d_nodes; host = f->h_nodes.data(); elems = f->h_nodes.size(); break;
        case SBGI_BUF_TRI_INDEX: dev = f->d_tri_index; host = f->h_tri_index.data(); elems = f->h_tri_index.size(); break;
        case SBGI_BUF_VERTEX_DATA: dev = f->d_vdata; host = f->h_vdata.data(); elems = f->h_vdata.size(); break;
        case SBGI_BUF_TRI_MATERIAL: dev = f->d_tri_mat; host = f->h_tri_mat.data(); elems = f->h_tri_mat.size(); break;
        case SBGI_BUF_MATERIALS: dev = f->d_mats; host = f->h_mats.data(); elems = f->h_mats.size(); break;
        case SBGI_BUF_LIGHTS: dev = f->d_lights; host = f->h_lights.data(); elems = f->h_lights.size(); break;
        default: return fail(f, SBGI_E_INVALID, "buffer id not updatable");
    }
    for (uint32_t r = 0; r < n_ranges; r++)
        if ((uint64_t)ranges[2 * r] + ranges[2 * r + 1] > elems) return fail(f, SBGI_E_INVALID, "update range out of bounds");
    HIP_OK(f, hipSetDevice(f->desc.device));
    // vertex positions and light records cannot break the layout: no validation, no host
    // sync (pageable uploads are staged before the call returns)
    const bool structural = buf_id != SBGI_BUF_VERTEX_DATA && buf_id != SBGI_BUF_LIGHTS;
    int32_t rc = SBGI_OK;
    std::vector<char> backup;
    if (structural) {
        if (buf_id == SBGI_BUF_BVH_NODES && (rc = sync_nodes(f)) != SBGI_OK) return rc;
        HIP_OK(f, hipStreamSynchronize(f->stream));
        backup.assign(static_cast<char *>(host), static_cast<char *>(host) + elems * esz);
    }
    for (uint32_t r = 0; r < n_ranges; r++) {
        uint64_t st = ranges[2 * r], cnt = ranges[2 * r + 1];
        memcpy(static_cast<char *>(host) + st * esz, static_cast<const char *>(src) + st * esz, cnt * esz);
    }
    if (structural && (rc = validate_scene(f)) != SBGI_OK) {
        memcpy(host, backup.data(), backup.size());
        return rc;
    }
    for (uint32_t r = 0; r < n_ranges; r++) {
        uint64_t st = ranges[2 * r], cnt = ranges[2 * r + 1];
        if (cnt)
            HIP_OK(f, hipMemcpyAsync(static_cast<char *>(dev) + st * esz, static_cast<char *>(host) + st * esz,
                                     cnt * esz, hipMemcpyHostToDevice, f->stream));
    }
    if (buf_id == SBGI_BUF_LIGHTS) {
        if ((rc = build_light_grid(f)) != SBGI_OK) return rc;
        return build_shadow_lists(f);
    }
    if ((rc = repack_all(f)) != SBGI_OK) return rc;
    if (buf_id == SBGI_BUF_BVH_NODES || buf_id == SBGI_BUF_MATERIALS) {
        if ((rc = sync_nodes(f)) != SBGI_OK) return rc;
        HIP_OK(f, hipStreamSynchronize(f->stream));
        // walk-format node copies depend on the pool and, in ordered mode, on the root set
        if ((rc = upload_walk_nodes(f)) != SBGI_OK) return rc;
        if (buf_id == SBGI_BUF_MATERIALS) {
            for (uint32_t i = 0; i < f->view.inst_count; i++)
                invert_affine_rows(f->h_mats.data() + f->view.inst_base + (size_t)i * 28,
                                   f->h_world.data() + (size_t)i * 16);
            HIP_OK(f, hipMemcpyAsync(f->d_world, f->h_world.data(), f->h_world.size() * sizeof(double),
                                     hipMemcpyHostToDevice, f->stream));
        }
        // instance bounds follow the new roots / records; the MT_ORDERED SAH TLAS is rebuilt
        // (a caller-supplied reference TLAS stands as uploaded)
        if ((rc = device_tlas(f, false)) != SBGI_OK) return rc;
    }
    return SBGI_OK;
}

int32_t sbgi_refit(sbgi_field *f, const uint32_t *blas_ranges, uint32_t n_ranges, int32_t refit_tlas) {
    if (!f || (n_ranges && !blas_ranges)) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    HIP_OK(f, hipSetDevice(f->desc.device));
    const sbgi_scene_view &v = f->view;
    // createFlatBlasRefitStepper validation (js/gi_refit.js:37-43, 95-139)
    for (uint32_t r = 0; r < n_ranges; r++) {
        uint32_t root = blas_ranges[2 * r], end = blas_ranges[2 * r + 1];
        if (!(end > root) || end > v.node_count) return fail(f, SBGI_E_LAYOUT, "refit range out of bounds");
        for (uint32_t i = root; i < end; i++) {
            const uint32_t *nd = f->h_nodes.data() + (size_t)i * 8;
            if (nd[7] == 0xFFFFFFFFu) {
                uint32_t l = i + 1;
                if (l >= end) return fail(f, SBGI_E_LAYOUT, "interior node without a left child in range");
                uint32_t rr = f->h_nodes[(size_t)l * 8 + 6];
                if (!(rr > l && rr < end)) return fail(f, SBGI_E_LAYOUT, "right child outside (left, end)");
            }
        }
    }
    int32_t rc = SBGI_OK;
    if (n_ranges && ((rc = record_kev(f, 0)) != SBGI_OK || (rc = record_kev(f, 1)) != SBGI_OK)) return rc;
    if (n_ranges) {
        if ((rc = build_plan(f, blas_ranges, n_ranges, f->call_plan)) != SBGI_OK) return rc;
        HIP_OK(f, launch_refit_levels(f->d_nodes, f->d_tri_index, f->d_vdata, f->call_plan.d_order,
                                      f->call_plan.off.data(), (uint32_t)f->call_plan.off.size() - 1, f->stream));
        HIP_OK(f, launch_oct_gather(f->d_walk_nodes, f->d_nodes, f->d_walk_perm, (ordered(f) ? 8u : 1u) * v.node_count,
                                    f->stream));
        f->nodes_dirty = true;
        if ((rc = record_kev(f, 2)) != SBGI_OK) return rc;
        f->kev_blas = true;
    }
    if (refit_tlas) {
        if ((rc = record_kev(f, 3)) != SBGI_OK || (rc = device_tlas(f, true)) != SBGI_OK ||
            (rc = record_kev(f, 4)) != SBGI_OK)
            return rc;
        f->kev_tlas = true;
    }
    return SBGI_OK;
}

int32_t sbgi_set_instance_transforms(sbgi_field *f, const uint32_t *slots, uint32_t n, const double *world,
                                     const uint8_t *active) {
    if (!f || (n && (!slots || !world))) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    const sbgi_scene_view &v = f->view;
    if (n > v.inst_count) return fail(f, SBGI_E_INVALID, "more slots than instances");
    for (uint32_t q = 0; q < n; q++)
        if (slots[q] >= v.inst_count) return fail(f, SBGI_E_INVALID, "instance slot out of range");
    HIP_OK(f, hipSetDevice(f->desc.device));
    // records on the host mirror, then one staged upload {worlds, records, slots} and a scatter
    void *stage = nullptr;
    int slot = 0;
    const size_t wb = (size_t)n * 16 * sizeof(double), rb = (size_t)n * 28 * sizeof(float);
    int32_t rc = stage_acquire(f, wb + rb + (size_t)n * sizeof(uint32_t), &stage, &slot);
    if (rc != SBGI_OK) return rc;
    double *sw = static_cast<double *>(stage);
    float *sr = reinterpret_cast<float *>(static_cast<char *>(stage) + wb);
    uint32_t *ss = reinterpret_cast<uint32_t *>(static_cast<char *>(stage) + wb + rb);
    for (uint32_t q = 0; q < n; q++) {
        const uint32_t s = slots[q];
        const double *e = world + (size_t)q * 16;
        memcpy(f->h_world.data() + (size_t)s * 16, e, 16 * sizeof(double));
        float *rec = f->h_mats.data() + v.inst_base + (size_t)s * 28;
        // inverse rows, det sign, active flag (js/spectral_scene.js:1285-1297, 1339-1351)
        double inv[16];
        sbgi::invert4(e, inv);
        static const int src[12] = {0, 4, 8, 12, 1, 5, 9, 13, 2, 6, 10, 14};  // column-major -> rows
        for (int i = 0; i < 12; i++) rec[i] = (float)inv[src[i]];
        rec[14] = sbgi::det4(e) < 0 ? -1.f : 1.f;
        if (active) rec[15] = active[q] ? 1.f : 0.f;
        memcpy(sw + (size_t)q * 16, e, 16 * sizeof(double));
        memcpy(sr + (size_t)q * 28, rec, 28 * sizeof(float));
        ss[q] = s;
    }
    if ((rc = record_kev(f, 3)) != SBGI_OK) return rc;
    if (n) {
        char *dst = static_cast<char *>(f->d_inst_stage);
        const size_t dwb = (size_t)v.inst_count * 16 * sizeof(double), drb = (size_t)v.inst_count * 28 * sizeof(float);
        HIP_OK(f, hipMemcpyAsync(dst, sw, wb, hipMemcpyHostToDevice, f->stream));
        HIP_OK(f, hipMemcpyAsync(dst + dwb, sr, rb, hipMemcpyHostToDevice, f->stream));
        HIP_OK(f, hipMemcpyAsync(dst + dwb + drb, ss, (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, f->stream));
        if ((rc = stage_commit(f, slot)) != SBGI_OK) return rc;
        HIP_OK(f, launch_scatter_inst(reinterpret_cast<const uint32_t *>(dst + dwb + drb),
                                      reinterpret_cast<const float *>(dst + dwb), reinterpret_cast<const double *>(dst),
                                      n, f->d_mats, v.inst_base, f->d_world, f->d_inst, f->stream));
    }
    if ((rc = device_tlas(f, true)) != SBGI_OK || (rc = record_kev(f, 4)) != SBGI_OK) return rc;
    f->kev_tlas = true;
    return SBGI_OK;
}

int32_t sbgi_set_skins(sbgi_field *f, const sbgi_skin_set *sets, uint32_t n_sets, uint32_t n_bones) {
    if (!f || (n_sets && !sets)) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    const sbgi_scene_view &v = f->view;
    if (n_sets && (n_bones == 0 || n_bones > 65536)) return fail(f, SBGI_E_INVALID, "bone count outside 1..65536");
    // validate: vertex ranges and BLAS ranges in bounds and disjoint, threaded BLAS layout,
    // joints inside the palette
    std::vector<std::pair<uint32_t, uint32_t>> vr, nr;
    uint64_t nv = 0;
    for (uint32_t s = 0; s < n_sets; s++) {
        const sbgi_skin_set &k = sets[s];
        if ((uint64_t)k.vertex_start + k.vertex_count > v.vertex_count || !k.vertex_count)
            return fail(f, SBGI_E_INVALID, "skin vertex range out of bounds");
        if (!(k.blas_end > k.blas_root) || k.blas_end > v.node_count)
            return fail(f, SBGI_E_LAYOUT, "skin BLAS range out of bounds");
        if (!k.joints || !k.weights) return fail(f, SBGI_E_INVALID, "skin set without joints or weights");
        for (uint64_t q = 0; q < 4ull * k.vertex_count; q++)
            if (k.joints[q] >= n_bones) return fail(f, SBGI_E_INVALID, "joint index outside the bone palette");
        for (uint32_t i = k.blas_root; i < k.blas_end; i++) {
            const uint32_t *nd = f->h_nodes.data() + (size_t)i * 8;
            if (nd[7] != 0xFFFFFFFFu) continue;
            const uint32_t l = i + 1;
            if (l >= k.blas_end) return fail(f, SBGI_E_LAYOUT, "interior node without a left child in range");
            const uint32_t rr = f->h_nodes[(size_t)l * 8 + 6];
            if (!(rr > l && rr < k.blas_end)) return fail(f, SBGI_E_LAYOUT, "right child outside (left, end)");
        }
        vr.emplace_back(k.vertex_start, k.vertex_start + k.vertex_count);
        nr.emplace_back(k.blas_root, k.blas_end);
        nv += k.vertex_count;
    }
    std::sort(vr.begin(), vr.end());
    std::sort(nr.begin(), nr.end());
    for (size_t i = 1; i < vr.size(); i++)
        if (vr[i].first < vr[i - 1].second || nr[i].first < nr[i - 1].second)
            return fail(f, SBGI_E_INVALID, "skin sets overlap");
    if (nv >= (1ull << 32)) return fail(f, SBGI_E_CAPACITY, "too many skinned vertices");
    HIP_OK(f, hipSetDevice(f->desc.device));
    HIP_OK(f, hipStreamSynchronize(f->stream));
    free_skins(f);
    if (!n_sets) return SBGI_OK;
    // rest pose = the vertexData last provided by the host
    std::vector<uint32_t> vtx, plan_ranges;
    std::vector<float4> rest, w;
    std::vector<uint2> joints, nranges, tranges;
    vtx.reserve(nv);
    rest.reserve(2 * nv);
    w.reserve(nv);
    joints.reserve(nv);
    uint32_t nmax = 0, tmax = 0;
    for (uint32_t s = 0; s < n_sets; s++) {
        const sbgi_skin_set &k = sets[s];
        for (uint32_t q = 0; q < k.vertex_count; q++) {
            const uint32_t pv = k.vertex_start + q;
            const float *d = f->h_vdata.data() + (size_t)pv * 8;
            vtx.push_back(pv);
            rest.push_back(make_float4(d[0], d[1], d[2], 0.f));
            rest.push_back(make_float4(d[3], d[4], d[5], 0.f));
            const uint16_t *j = k.joints + (size_t)q * 4;
            joints.push_back(make_uint2(j[0] | ((uint32_t)j[1] << 16), j[2] | ((uint32_t)j[3] << 16)));
            const float *ww = k.weights + (size_t)q * 4;
            w.push_back(make_float4(ww[0], ww[1], ww[2], ww[3]));
        }
        uint32_t t0 = UINT32_MAX, t1 = 0;
        for (uint32_t i = k.blas_root; i < k.blas_end; i++) {
            const uint32_t info = f->h_nodes[(size_t)i * 8 + 7];
            if (info == 0xFFFFFFFFu) continue;
            const uint32_t off = info & 0x00FFFFFFu, cnt = info >> 24;
            if (cnt) {
                t0 = std::min(t0, off);
                t1 = std::max(t1, off + cnt);
            }
        }
        nranges.push_back(make_uint2(k.blas_root, k.blas_end));
        plan_ranges.push_back(k.blas_root);
        plan_ranges.push_back(k.blas_end);
        if (t1 > t0) tranges.push_back(make_uint2(t0, t1));
        nmax = std::max(nmax, k.blas_end - k.blas_root);
        if (t1 > t0) tmax = std::max(tmax, t1 - t0);
    }
    if (dalloc(&f->d_skin_vtx, vtx.size() * 4) != hipSuccess || dalloc(&f->d_skin_rest, rest.size() * 16) != hipSuccess ||
        dalloc(&f->d_skin_joints, joints.size() * 8) != hipSuccess || dalloc(&f->d_skin_w, w.size() * 16) != hipSuccess ||
        dalloc(&f->d_bones, (size_t)n_bones * 48) != hipSuccess ||
        dalloc(&f->d_skin_nranges, nranges.size() * 8) != hipSuccess ||
        dalloc(&f->d_skin_tranges, std::max<size_t>(1, tranges.size()) * 8) != hipSuccess) {
        free_skins(f);
        return fail(f, SBGI_E_NOMEM, "skin allocation");
    }
    HIP_OK(f, hipMemcpy(f->d_skin_vtx, vtx.data(), vtx.size() * 4, hipMemcpyHostToDevice));
    HIP_OK(f, hipMemcpy(f->d_skin_rest, rest.data(), rest.size() * 16, hipMemcpyHostToDevice));
    HIP_OK(f, hipMemcpy(f->d_skin_joints, joints.data(), joints.size() * 8, hipMemcpyHostToDevice));
    HIP_OK(f, hipMemcpy(f->d_skin_w, w.data(), w.size() * 16, hipMemcpyHostToDevice));
    HIP_OK(f, hipMemcpy(f->d_skin_nranges, nranges.data(), nranges.size() * 8, hipMemcpyHostToDevice));
    if (!tranges.empty()) HIP_OK(f, hipMemcpy(f->d_skin_tranges, tranges.data(), tranges.size() * 8, hipMemcpyHostToDevice));
    f->skin_vtx_n = (uint32_t)vtx.size();
    f->skin_bones_n = n_bones;
    f->skin_ranges_n = (uint32_t)nranges.size();
    f->skin_tranges_n = (uint32_t)tranges.size();
    f->skin_nmax = nmax;
    f->skin_tmax = tmax;
    int32_t rc = build_plan(f, plan_ranges.data(), n_sets, f->skin_plan);
    if (rc != SBGI_OK) {
        free_skins(f);
        return rc;
    }
    HIP_OK(f, hipStreamSynchronize(f->stream));
    return SBGI_OK;
}

int32_t sbgi_skin(sbgi_field *f, const float *bones, uint32_t n_bones, int32_t update_tlas) {
    if (!f || !bones) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    if (!f->skin_vtx_n) return fail(f, SBGI_E_STATE, "no skin sets registered");
    if (n_bones != f->skin_bones_n) return fail(f, SBGI_E_INVALID, "bone count differs from the registered palette");
    HIP_OK(f, hipSetDevice(f->desc.device));
    void *stage = nullptr;
    int slot = 0;
    const size_t bytes = (size_t)n_bones * 48;
    int32_t rc = stage_acquire(f, bytes, &stage, &slot);
    if (rc != SBGI_OK) return rc;
    memcpy(stage, bones, bytes);
    const sbgi_scene_view &v = f->view;
    if ((rc = record_kev(f, 0)) != SBGI_OK) return rc;
    HIP_OK(f, hipMemcpyAsync(f->d_bones, stage, bytes, hipMemcpyHostToDevice, f->stream));
    if ((rc = stage_commit(f, slot)) != SBGI_OK) return rc;
    HIP_OK(f, launch_skin(f->d_skin_vtx, f->d_skin_rest, f->d_skin_joints, f->d_skin_w, f->d_bones, f->skin_vtx_n,
                          f->d_vdata, f->stream));
    if ((rc = record_kev(f, 1)) != SBGI_OK) return rc;
    HIP_OK(f, launch_refit_levels(f->d_nodes, f->d_tri_index, f->d_vdata, f->skin_plan.d_order,
                                  f->skin_plan.off.data(), (uint32_t)f->skin_plan.off.size() - 1, f->stream));
    HIP_OK(f, launch_oct_gather_ranges(f->d_walk_nodes, f->d_nodes, f->d_walk_perm, v.node_count, ordered(f) ? 8u : 1u,
                                       f->d_skin_nranges, f->skin_ranges_n, f->skin_nmax, f->stream));
    HIP_OK(f, launch_pack_tris_ranges(f->d_tri_index, f->d_vdata, f->d_tri_mat, f->d_mats, v.tri_count,
                                      f->d_skin_tranges, f->skin_tranges_n, f->skin_tmax, f->d_tri, f->stream));
    if ((rc = record_kev(f, 2)) != SBGI_OK) return rc;
    f->nodes_dirty = true;
    f->kev_blas = true;
    if (update_tlas) {
        if ((rc = record_kev(f, 3)) != SBGI_OK || (rc = device_tlas(f, true)) != SBGI_OK ||
            (rc = record_kev(f, 4)) != SBGI_OK)
            return rc;
        f->kev_tlas = true;
    }
    return SBGI_OK;
}

// Every by-value kernel argument and launch choice of one tick: a captured tick graph may
// be replayed only while this is unchanged, byte for byte.
static void tick_key(const sbgi_field *f, const CascadeState &cs, const DevScene &s, const DevCascade &c,
                     const sbgi_tick_params &p, std::vector<unsigned char> &key) {
    key.clear();
    auto put = [&](const void *v, size_t n) {
        const unsigned char *b = static_cast<const unsigned char *>(v);
        key.insert(key.end(), b, b + n);
    };
    put(&s, sizeof(s));
    put(&c, sizeof(c));
    put(&p, sizeof(p));
    const void *ptrs[] = {f->dirs, f->rays, f->hit_ids, f->counters, f->tw.hits, f->tw.vis, f->tw.queue.rec,
                          f->tw.queue.count, cs.irr, cs.depth, cs.meta, cs.atlas, cs.datlas, cs.evis, cs.rough,
                          cs.gnum, cs.gden, cs.ratlas, cs.gatlas, f->rtables};
    put(ptrs, sizeof(ptrs));
    const uint32_t words[] = {f->tw.queue.capacity, f->count_mode ? 1u : 0u, f->desc.blend_precision,
                              f->desc.reflections, f->stage_timing ? 1u : 0u};
    put(words, sizeof(words));
}

static void drop_tick_graph(sbgi_field *f) {
    if (f->tick_exec) (void)hipGraphExecDestroy(f->tick_exec);
    if (f->tick_graph) (void)hipGraphDestroy(f->tick_graph);
    f->tick_exec = nullptr;
    f->tick_graph = nullptr;
    f->tick_key.clear();
}

// The ordered solve of one tick (trace -> emitter visibility -> reflections + blend ->
// upload) with its stage events, marked through mk.
static int32_t enqueue_tick(sbgi_field *f, CascadeState &cs, const DevScene &s, const DevCascade &c,
                            const sbgi_tick_params *p, hipStream_t st, const StageMarker &mk) {
    HIP_OK(f, mk(f->ev[0], st));
    if (p->needs_clear) {
        HIP_OK(f, launch_clear_atlas(cs.atlas, (size_t)cs.aw * cs.ah, st));
        HIP_OK(f, launch_clear_atlas(cs.datlas, (size_t)cs.daw * cs.dah, st));
        HIP_OK(f, launch_fill_u32(reinterpret_cast<uint32_t *>(cs.evis), 0u, cs.total * SBGI_MAX_EMITTERS, st));
        HIP_OK(f, clear_refl_atlases(cs, st));
    }
    HIP_OK(f, launch_raydirs(f->dirs, c.rays, p->frame_jitter, st));
    HIP_OK(f, launch_weights(c, *p, f->dirs, f->bt, f->desc.blend_precision, st));
    if (f->rtables) HIP_OK(f, launch_weights_refl(c.rays, f->dirs, f->rt, f->desc.blend_precision, st));
    HIP_OK(f, mk(f->ev[1], st));
    HIP_OK(f, launch_trace(s, c, *p, f->dirs, cs.meta, cs.atlas, nullptr, f->tw, f->rays, f->hit_ids,
                           f->count_mode ? f->counters : nullptr, f->tev, st, mk));
    HIP_OK(f, mk(f->eev, st));
    // the emitter rays' table takes the hit records, which the shading has consumed
    HIP_OK(f, launch_emitter_vis(s, c, *p, cs.meta, cs.evis, f->tw.hits, f->tw.queue.count + trace_heads(3), st));
    HIP_OK(f, mk(f->ev[2], st));
    HIP_OK(f, launch_reflect_and_blend(f, cs, c, *p, f->rays, st));
    HIP_OK(f, mk(f->ev[3], st));
    HIP_OK(f, launch_upload(c, *p, cs.irr, cs.depth, cs.atlas, cs.datlas, nullptr, st, cs.rough, cs.ratlas));
    HIP_OK(f, mk(f->ev[4], st));
    return SBGI_OK;
}

// Builds a tick graph as captured pieces separated by event-record nodes: each stage
// boundary closes the piece captured so far and starts the next one. (Event records made
// during a stream capture are not graph nodes on every HIP runtime; explicit event nodes
// between child graphs are, so replays time their stages like eager ticks.)
struct TickGraphBuild {
    std::vector<hipGraph_t> pieces;
    std::vector<hipEvent_t> events;  // events[i] follows pieces[i]
    static hipError_t mark(void *ctx, hipEvent_t e, hipStream_t st) {
        TickGraphBuild *b = static_cast<TickGraphBuild *>(ctx);
        hipGraph_t g = nullptr;
        hipError_t err = hipStreamEndCapture(st, &g);
        if (err != hipSuccess) return err;
        b->pieces.push_back(g);
        b->events.push_back(e);
        return hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    }
    ~TickGraphBuild() {
        for (hipGraph_t g : pieces)
            if (g) (void)hipGraphDestroy(g);
    }
};

// Without stage timing the tick is captured as one piece: no stage boundary is marked.
static hipError_t no_mark(void *, hipEvent_t, hipStream_t) { return hipSuccess; }

static hipError_t build_tick_graph(sbgi_field *f, CascadeState &cs, const DevScene &s, const DevCascade &c,
                                   const sbgi_tick_params *p, hipStream_t st, hipGraph_t *out) {
    *out = nullptr;
    TickGraphBuild b;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const StageMarker mk = f->stage_timing ? StageMarker{&TickGraphBuild::mark, &b} : StageMarker{&no_mark, nullptr};
    const int32_t rc = enqueue_tick(f, cs, s, c, p, st, mk);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(st, &cap);
    hipGraph_t last = nullptr;
    if (cap != hipStreamCaptureStatusNone) {
        const hipError_t ee = hipStreamEndCapture(st, &last);
        if (e == hipSuccess) e = ee;
    }
    b.pieces.push_back(last);
    if (rc != SBGI_OK) return hipErrorUnknown;
    if (e != hipSuccess) return e;
    hipGraph_t g = nullptr;
    if ((e = hipGraphCreate(&g, 0)) != hipSuccess) return e;
    hipGraphNode_t prev = nullptr;
    for (size_t i = 0; i < b.pieces.size() && e == hipSuccess; i++) {
        size_t nodes = 0;
        if (b.pieces[i] && hipGraphGetNodes(b.pieces[i], nullptr, &nodes) == hipSuccess && nodes) {
            hipGraphNode_t n = nullptr;
            e = hipGraphAddChildGraphNode(&n, g, prev ? &prev : nullptr, prev ? 1 : 0, b.pieces[i]);
            prev = n;
        }
        if (e == hipSuccess && i < b.events.size()) {
            hipGraphNode_t n = nullptr;
            e = hipGraphAddEventRecordNode(&n, g, prev ? &prev : nullptr, prev ? 1 : 0, b.events[i]);
            prev = n;
        }
    }
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return e;
    }
    *out = g;
    return hipSuccess;
}

int32_t sbgi_tick(sbgi_field *f, const sbgi_tick_params *p, uint64_t *ticket) {
    if (!f || !p) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene: call sbgi_set_scene first");
    if (p->cascade >= f->desc.cascade_count) return fail(f, SBGI_E_INVALID, "cascade index out of range");
    CascadeState &cs = f->casc[p->cascade];
    const uint32_t updated = p->updated_count ? std::min(p->updated_count, cs.total) : cs.total;
    if (p->probe_offset >= cs.total) return fail(f, SBGI_E_INVALID, "probe_offset out of range");
    HIP_OK(f, hipSetDevice(f->desc.device));
    const DevCascade c = make_cascade(f, cs, p->probe_offset, updated);
    const DevScene s = make_scene(f);
    hipStream_t st = f->stream;
    f->diff_valid = false;
    // One hipGraph per frame: a tick whose kernel arguments equal the captured tick's
    // replays it; another is captured and launched (the first tick, needs_clear ticks and
    // fields created with SBGI_FIELD_NO_GRAPH run eagerly). Device data (the scene, the
    // temporal state) is read at replay time, so edits between ticks need no recapture.
    const bool graph = !(f->desc.flags & SBGI_FIELD_NO_GRAPH) && !p->needs_clear && !f->tick_graph_failed;
    bool replayed = false, eager = !graph;
    if (graph) {
        tick_key(f, cs, s, c, *p, f->scratch_key);
        if (f->tick_exec && f->scratch_key == f->tick_key) {
            HIP_OK(f, hipGraphLaunch(f->tick_exec, st));
            f->tick_replays++;
            replayed = true;
        } else {
            drop_tick_graph(f);
            hipGraph_t g = nullptr;
            hipGraphExec_t x = nullptr;
            hipError_t e = build_tick_graph(f, cs, s, c, p, st, &g);
            if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            if (e == hipSuccess) {
                f->tick_graph = g;
                f->tick_exec = x;
                f->tick_key = f->scratch_key;
                HIP_OK(f, hipGraphLaunch(f->tick_exec, st));
            } else {
                // graphs not available here: this field ticks eagerly from now on
                if (g) (void)hipGraphDestroy(g);
                (void)hipGetLastError();
                f->tick_graph_failed = true;
                eager = true;
            }
        }
    }
    if (eager) {
        const int32_t rc = enqueue_tick(f, cs, s, c, p, st, StageMarker{});
        if (rc != SBGI_OK) return rc;
    }
    f->ticket++;
    HIP_OK(f, hipEventRecord(f->ring[f->ticket % kTicketRing], st));
    f->last_updated = updated;
    f->last_cascade = p->cascade;
    f->last_replayed = replayed;
    f->last_timed = eager || f->stage_timing;
    f->stats_valid = true;
    if (ticket) *ticket = f->ticket;
    return SBGI_OK;
}

int32_t sbgi_poll(sbgi_field *f, uint64_t ticket) {
    if (!f) return SBGI_E_INVALID;
    if (ticket > f->ticket) return fail(f, SBGI_E_INVALID, "unknown ticket");
    if (ticket == 0) return 1;
    // A ticket older than the ring is covered by the oldest event still in it: that event
    // was recorded later on the same in-order stream, so once it has completed the older
    // ticket has too. Until then the older ticket reads as pending.
    const uint64_t slot = (f->ticket - ticket < kTicketRing) ? ticket : f->ticket + 1u;
    hipError_t q = hipEventQuery(f->ring[slot % kTicketRing]);
    if (q == hipSuccess) return 1;
    if (q == hipErrorNotReady) return 0;
    return hip_fail(f, q, "hipEventQuery");
}

int32_t sbgi_wait(sbgi_field *f, uint64_t ticket, int32_t timeout_ms) {
    if (!f) return SBGI_E_INVALID;
    if (ticket > f->ticket) return fail(f, SBGI_E_INVALID, "unknown ticket");
    if (ticket == 0) return SBGI_OK;
    hipEvent_t e = (f->ticket - ticket < kTicketRing) ? f->ring[ticket % kTicketRing] : f->ring[f->ticket % kTicketRing];
    if (timeout_ms < 0) {
        HIP_OK(f, hipEventSynchronize(e));
        return SBGI_OK;
    }
    auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        hipError_t q = hipEventQuery(e);
        if (q == hipSuccess) return SBGI_OK;
        if (q != hipErrorNotReady) return hip_fail(f, q, "hipEventQuery");
        auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
        if (ms >= timeout_ms) return fail(f, SBGI_E_TIMEOUT, "sbgi_wait timed out");
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

size_t sbgi_buffer_bytes(const sbgi_field *f, int32_t buf_id, uint32_t cascade) {
    if (!f) return 0;
    const uint32_t R = f->desc.rays_per_probe;
    if (buf_id >= SBGI_BUF_IRR_STATE && buf_id <= SBGI_BUF_DEPTH_ATLAS) {
        if (cascade >= f->desc.cascade_count) return 0;
        const CascadeState &cs = f->casc[cascade];
        const size_t dt = f->desc.depth_oct_res == 6 ? 64 : (size_t)cs.tile_d * cs.tile_d;
        switch (buf_id) {
            case SBGI_BUF_IRR_STATE: return (size_t)cs.total * 64 * 16;
            case SBGI_BUF_DEPTH_STATE: return (size_t)cs.total * dt * 8;
            case SBGI_BUF_META: return (size_t)cs.total * 16;
            case SBGI_BUF_IRR_ATLAS: return (size_t)cs.aw * cs.ah * 8;
            case SBGI_BUF_DEPTH_ATLAS: return (size_t)cs.daw * cs.dah * 8;
        }
    }
    switch (buf_id) {
        case SBGI_BUF_RAY_DATA: return (size_t)f->last_updated * R * 16;
        case SBGI_BUF_HIT_IDS: return f->hit_ids ? (size_t)f->last_updated * R * 8 : 0;
        case SBGI_BUF_SHADOW_VIS: return (size_t)f->last_updated * R * 8;
        case SBGI_BUF_MATERIALS: return f->h_mats.size() * 4;
        case SBGI_BUF_LIGHTS: return f->h_lights.size() * 4;
        case SBGI_BUF_NODES_REFIT:
        case SBGI_BUF_BVH_NODES: return f->h_nodes.size() * 4;
        case SBGI_BUF_TRI_PACKED: return f->has_scene ? (size_t)f->view.tri_count * 48 : 0;
        case SBGI_BUF_VERTEX_DATA: return f->h_vdata.size() * 4;
        case SBGI_BUF_TRI_INDEX: return f->h_tri_index.size() * 4;
        case SBGI_BUF_TRI_MATERIAL: return f->h_tri_mat.size() * 4;
        case SBGI_BUF_INST_BOUNDS: return f->has_scene ? (size_t)f->view.inst_count * 24 : 0;
        case SBGI_BUF_TLAS_SAH: return f->has_scene ? (size_t)f->tlas_sah_count * 32 : 0;
        case SBGI_BUF_TLAS_OCT: return f->has_scene && f->d_tlas_oct ? (size_t)f->tlas_sah_count * 8 * 32 : 0;
        case SBGI_BUF_EMITTER_VIS:
            return cascade < f->desc.cascade_count ? (size_t)f->casc[cascade].total * SBGI_MAX_EMITTERS * 4 : 0;
        case SBGI_BUF_LIGHT_GRID: return light_grid_cells(f) * (SBGI_LIGHTS_PER_CELL + 1) * 4;
        case SBGI_BUF_SHADOW_PROJ: return f->sl.list ? (size_t)f->view.light_count * kShadowProj * 16 : 0;
        case SBGI_BUF_SHADOW_CELLS: return f->sl.list ? (size_t)f->view.light_count * kShadowRecs * 8 : 0;
        case SBGI_BUF_SHADOW_LIST: return f->sl.list ? (size_t)f->sl.capacity * 2 : 0;
    }
    if (buf_id >= SBGI_BUF_ROUGH_STATE && buf_id <= SBGI_BUF_GLOSSY_ATLAS) {
        if (cascade >= f->desc.cascade_count || !f->casc[cascade].rough) return 0;
        const CascadeState &cs = f->casc[cascade];
        const size_t G = glossy_tile(f->desc.reflections);
        switch (buf_id) {
            case SBGI_BUF_ROUGH_STATE: return (size_t)cs.total * 64 * 16;
            case SBGI_BUF_ROUGH_ATLAS: return (size_t)cs.aw * cs.ah * 8;
            case SBGI_BUF_GLOSSY_STATE: return (size_t)cs.total * G * G * 16;
            case SBGI_BUF_GLOSSY_WEIGHT: return (size_t)cs.total * G * G * 4;
            case SBGI_BUF_GLOSSY_ATLAS: return (size_t)cs.gaw * cs.gah * 8;
        }
    }
    return 0;
}

int32_t sbgi_read(sbgi_field *f, int32_t buf_id, uint32_t cascade, void *dst, size_t bytes) {
    if (!f || !dst) return SBGI_E_INVALID;
    const size_t need = sbgi_buffer_bytes(f, buf_id, cascade);
    if (need == 0 || bytes != need) return fail(f, SBGI_E_INVALID, "buffer id/cascade/size mismatch");
    HIP_OK(f, hipSetDevice(f->desc.device));
    HIP_OK(f, hipStreamSynchronize(f->stream));
    const void *src = nullptr;
    if (buf_id >= SBGI_BUF_IRR_STATE && buf_id <= SBGI_BUF_DEPTH_ATLAS) {
        const CascadeState &cs = f->casc[cascade];
        src = buf_id == SBGI_BUF_IRR_STATE     ? (const void *)cs.irr
              : buf_id == SBGI_BUF_DEPTH_STATE ? (const void *)cs.depth
              : buf_id == SBGI_BUF_META        ? (const void *)cs.meta
              : buf_id == SBGI_BUF_IRR_ATLAS   ? (const void *)cs.atlas
                                               : (const void *)cs.datlas;
        HIP_OK(f, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return SBGI_OK;
    }
    if (buf_id == SBGI_BUF_RAY_DATA || buf_id == SBGI_BUF_HIT_IDS || buf_id == SBGI_BUF_SHADOW_VIS) {
        // device order is k-major (k * updated + slot); the reference order is slot * R + k
        const uint32_t R = f->desc.rays_per_probe, U = f->last_updated;
        const size_t es = buf_id == SBGI_BUF_RAY_DATA ? 16 : 8;
        std::vector<char> tmp(bytes);
        src = buf_id == SBGI_BUF_RAY_DATA  ? (const void *)f->rays
              : buf_id == SBGI_BUF_HIT_IDS ? (const void *)f->hit_ids
                                           : (const void *)f->tw.vis;
        HIP_OK(f, hipMemcpy(tmp.data(), src, bytes, hipMemcpyDeviceToHost));
        char *o = static_cast<char *>(dst);
        for (uint32_t k = 0; k < R; k++)
            for (uint32_t sl = 0; sl < U; sl++) memcpy(o + ((size_t)sl * R + k) * es, tmp.data() + ((size_t)k * U + sl) * es, es);
        return SBGI_OK;
    }
    if (buf_id == SBGI_BUF_MATERIALS) {
        HIP_OK(f, hipMemcpy(dst, f->d_mats, bytes, hipMemcpyDeviceToHost));
        return SBGI_OK;
    }
    if (buf_id == SBGI_BUF_LIGHTS) {
        HIP_OK(f, hipMemcpy(dst, f->d_lights, bytes, hipMemcpyDeviceToHost));
        return SBGI_OK;
    }
    if (buf_id == SBGI_BUF_TRI_PACKED || buf_id == SBGI_BUF_VERTEX_DATA || buf_id == SBGI_BUF_INST_BOUNDS ||
        buf_id == SBGI_BUF_TLAS_SAH || buf_id == SBGI_BUF_TLAS_OCT || buf_id == SBGI_BUF_TRI_INDEX ||
        buf_id == SBGI_BUF_TRI_MATERIAL) {
        src = buf_id == SBGI_BUF_TRI_PACKED     ? (const void *)f->d_tri
              : buf_id == SBGI_BUF_VERTEX_DATA  ? (const void *)f->d_vdata
              : buf_id == SBGI_BUF_INST_BOUNDS  ? (const void *)f->d_inst_bounds
              : buf_id == SBGI_BUF_TRI_INDEX    ? (const void *)f->d_tri_index
              : buf_id == SBGI_BUF_TRI_MATERIAL ? (const void *)f->d_tri_mat
              : buf_id == SBGI_BUF_TLAS_OCT     ? (const void *)f->d_tlas_oct
                                                : (const void *)f->d_tlas_sah;
        HIP_OK(f, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return SBGI_OK;
    }
    if (buf_id == SBGI_BUF_SHADOW_PROJ || buf_id == SBGI_BUF_SHADOW_CELLS || buf_id == SBGI_BUF_SHADOW_LIST) {
        src = buf_id == SBGI_BUF_SHADOW_PROJ    ? (const void *)f->sl.proj
              : buf_id == SBGI_BUF_SHADOW_CELLS ? (const void *)f->sl.cells
                                                : (const void *)f->sl.list;
        HIP_OK(f, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return SBGI_OK;
    }
    if (buf_id == SBGI_BUF_BVH_NODES || buf_id == SBGI_BUF_NODES_REFIT) {
        HIP_OK(f, hipMemcpy(dst, f->d_nodes, bytes, hipMemcpyDeviceToHost));
        return SBGI_OK;
    }
    if (buf_id >= SBGI_BUF_ROUGH_STATE && buf_id <= SBGI_BUF_GLOSSY_ATLAS) {
        const CascadeState &cs = f->casc[cascade];
        src = buf_id == SBGI_BUF_ROUGH_STATE    ? (const void *)cs.rough
              : buf_id == SBGI_BUF_ROUGH_ATLAS  ? (const void *)cs.ratlas
              : buf_id == SBGI_BUF_GLOSSY_STATE ? (const void *)cs.gnum
              : buf_id == SBGI_BUF_GLOSSY_WEIGHT ? (const void *)cs.gden
                                                 : (const void *)cs.gatlas;
        if (!src) return fail(f, SBGI_E_INVALID, "buffer not present at this reflection quality");
        HIP_OK(f, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return SBGI_OK;
    }
    if (buf_id == SBGI_BUF_EMITTER_VIS || buf_id == SBGI_BUF_LIGHT_GRID) {
        src = buf_id == SBGI_BUF_EMITTER_VIS ? (const void *)f->casc[cascade].evis : (const void *)f->d_light_grid;
        HIP_OK(f, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return SBGI_OK;
    }
    return fail(f, SBGI_E_INVALID, "buffer not readable");
}

int32_t sbgi_write(sbgi_field *f, int32_t buf_id, uint32_t cascade, const void *src, size_t bytes) {
    if (!f || !src) return SBGI_E_INVALID;
    if (buf_id != SBGI_BUF_IRR_STATE && buf_id != SBGI_BUF_DEPTH_STATE && buf_id != SBGI_BUF_META &&
        buf_id != SBGI_BUF_EMITTER_VIS && buf_id != SBGI_BUF_ROUGH_STATE && buf_id != SBGI_BUF_GLOSSY_STATE &&
        buf_id != SBGI_BUF_GLOSSY_WEIGHT)
        return fail(f, SBGI_E_INVALID, "only the per-probe state buffers are writable");
    const size_t need = sbgi_buffer_bytes(f, buf_id, cascade);
    if (need == 0 || bytes != need) return fail(f, SBGI_E_INVALID, "buffer id/cascade/size mismatch");
    HIP_OK(f, hipSetDevice(f->desc.device));
    HIP_OK(f, hipStreamSynchronize(f->stream));
    const CascadeState &cs = f->casc[cascade];
    void *dst = buf_id == SBGI_BUF_IRR_STATE      ? (void *)cs.irr
                : buf_id == SBGI_BUF_DEPTH_STATE  ? (void *)cs.depth
                : buf_id == SBGI_BUF_META         ? (void *)cs.meta
                : buf_id == SBGI_BUF_ROUGH_STATE  ? (void *)cs.rough
                : buf_id == SBGI_BUF_GLOSSY_STATE ? (void *)cs.gnum
                : buf_id == SBGI_BUF_GLOSSY_WEIGHT ? (void *)cs.gden
                                                  : (void *)cs.evis;
    HIP_OK(f, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SBGI_OK;
}

int32_t sbgi_get_stats(sbgi_field *f, sbgi_stats *out) {
    if (!f || !out) return SBGI_E_INVALID;
    memset(out, 0, sizeof(*out));
    if (!f->stats_valid) return fail(f, SBGI_E_STATE, "no tick yet");
    if (f->last_timed) {
        HIP_OK(f, hipEventSynchronize(f->ev[4]));
        float ms[4];
        for (int i = 0; i < 4; i++) HIP_OK(f, hipEventElapsedTime(&ms[i], f->ev[i], f->ev[i + 1]));
        out->ms_raydir = ms[0];
        out->ms_trace = ms[1];
        out->ms_blend = ms[2];
        out->ms_upload = ms[3];
        HIP_OK(f, hipEventElapsedTime(&out->ms_total, f->ev[0], f->ev[4]));
    } else {
        HIP_OK(f, hipStreamSynchronize(f->stream));
    }
    out->updated = f->last_updated;
    out->rays = f->last_updated * f->desc.rays_per_probe;
    Counters c{};
    HIP_OK(f, hipMemcpy(&c, f->counters, sizeof(c), hipMemcpyDeviceToHost));
    out->blas_nodes = c.walk[0] + c.shadow_walk[0];
    out->tlas_nodes = c.walk[1] + c.shadow_walk[1];
    out->tri_tests = c.walk[2] + c.shadow_walk[2];
    out->inst_visits = c.walk[3] + c.shadow_walk[3];
    out->shadow_rays = c.shadow_rays;
    out->shadow_blas_nodes = c.shadow_walk[0];
    out->shadow_tlas_nodes = c.shadow_walk[1];
    out->shadow_tri_tests = c.shadow_walk[2];
    out->shadow_inst_visits = c.shadow_walk[3];
    for (int i = 0; i < 4; i++) {
        out->walk_waves[i] = c.walk[4 + i];
        out->shadow_waves[i] = c.shadow_walk[4 + i];
    }
    float sub[4] = {0.f, 0.f, 0.f, 0.f};
    if (f->last_updated && f->last_timed) {
        for (int i = 0; i < 3; i++) HIP_OK(f, hipEventElapsedTime(&sub[i], f->tev[i], f->tev[i + 1]));
        HIP_OK(f, hipEventElapsedTime(&sub[3], f->tev[3], f->eev));
        HIP_OK(f, hipEventElapsedTime(&out->ms_emitter_vis, f->eev, f->ev[2]));
    }
    out->ms_walk_primary = sub[0];
    out->ms_emit = sub[1];
    out->ms_walk_shadow = sub[2];
    out->ms_resolve = sub[3];
    out->tick_replayed = f->last_replayed ? 1u : 0u;
    out->stage_timed = f->last_timed ? 1u : 0u;
    return SBGI_OK;
}

int32_t sbgi_rebuild_blas(sbgi_field *f, const uint32_t *blas_ranges, uint32_t n_ranges) {
    if (!f || (n_ranges && !blas_ranges)) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    if (!n_ranges) return SBGI_OK;
    const sbgi_scene_view &v = f->view;
    // validate: threaded layout, odd node count, one contiguous triangle span per BLAS,
    // disjoint ranges, at most 255 triangles per leaf
    std::vector<uint4> desc(n_ranges);
    std::vector<uint32_t> prefix(n_ranges + 1, 0u);
    std::vector<std::pair<uint32_t, uint32_t>> nr, tr;
    uint64_t total = 0, nodes_total = 0;
    for (uint32_t r = 0; r < n_ranges; r++) {
        const uint32_t root = blas_ranges[2 * r], end = blas_ranges[2 * r + 1];
        if (!(end > root) || end > v.node_count) return fail(f, SBGI_E_LAYOUT, "rebuild range out of bounds");
        const uint32_t N = end - root;
        if (!(N & 1u)) return fail(f, SBGI_E_LAYOUT, "rebuild range with an even node count");
        uint32_t t0 = UINT32_MAX, t1 = 0;
        uint64_t T = 0;
        for (uint32_t i = root; i < end; i++) {
            const uint32_t *nd = f->h_nodes.data() + (size_t)i * 8;
            if (nd[7] == 0xFFFFFFFFu) {
                const uint32_t l = i + 1;
                if (l >= end) return fail(f, SBGI_E_LAYOUT, "interior node without a left child in range");
                const uint32_t rr = f->h_nodes[(size_t)l * 8 + 6];
                if (!(rr > l && rr < end)) return fail(f, SBGI_E_LAYOUT, "right child outside (left, end)");
            } else {
                const uint32_t off = nd[7] & 0x00FFFFFFu, cnt = nd[7] >> 24;
                if (!cnt) continue;
                t0 = std::min(t0, off);
                t1 = std::max(t1, off + cnt);
                T += cnt;
            }
        }
        const uint32_t L = (N + 1) / 2;
        if (!T || t1 - t0 != T) return fail(f, SBGI_E_LAYOUT, "BLAS leaves do not cover one contiguous triangle span");
        if (T < L || (T + L - 1) / L > 255) return fail(f, SBGI_E_LAYOUT, "BLAS triangle count does not fit its node count");
        desc[r] = make_uint4(root, N, t0, (uint32_t)T);
        prefix[r + 1] = prefix[r] + (uint32_t)T;
        nr.emplace_back(root, end);
        tr.emplace_back(t0, t1);
        total += T;
        nodes_total += N;
    }
    std::sort(nr.begin(), nr.end());
    std::sort(tr.begin(), tr.end());
    for (size_t i = 1; i < nr.size(); i++)
        if (nr[i].first < nr[i - 1].second || tr[i].first < tr[i - 1].second)
            return fail(f, SBGI_E_INVALID, "rebuild ranges overlap");
    if (total >= (1ull << 31)) return fail(f, SBGI_E_CAPACITY, "too many triangles to rebuild at once");
    HIP_OK(f, hipSetDevice(f->desc.device));
    hipStream_t st = f->stream;
    const uint32_t n = (uint32_t)total, copies = ordered(f) ? 8u : 1u;
    uint32_t range_bits = 0;
    while ((1u << range_bits) < n_ranges) range_bits++;
    const size_t sort_bytes = rebuild_sort_bytes(n);
    // depth bound of the top-down build and each depth's grid (nodes at depth d: at most
    // n_ranges 2^d, and at most the leaf count)
    uint32_t max_leaves = 0;
    for (uint32_t r = 0; r < n_ranges; r++) max_leaves = std::max(max_leaves, (desc[r].y + 1) / 2);
    const uint32_t depths = rebuild_depth_bound(max_leaves), cap = (uint32_t)nodes_total;
    const uint64_t leaves_total = (nodes_total + n_ranges) / 2;
    std::vector<uint32_t> blocks(depths);
    for (uint32_t d = 0; d < depths; d++) {
        const uint64_t at_d = d < 32 ? std::min<uint64_t>((uint64_t)n_ranges << d, leaves_total) : leaves_total;
        blocks[d] = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((at_d + 255) / 256, 1024));
    }
    // scratch, kept on the field and grown on demand (stream order separates rebuilds)
    const size_t sz_desc = (size_t)n_ranges * 16, sz_pref = ((size_t)n_ranges + 1) * 4, sz_cen = (size_t)n * 16,
                 sz_cb = (size_t)n_ranges * 24, sz_keys = (size_t)n * 8, sz_vals = (size_t)n * 4,
                 sz_old = (size_t)n * 16, sz_tasks = nodes_total * 16, sz_lv = ((size_t)depths * 2 + 3) * 4,
                 sz_order = nodes_total * 4, sz_q = ordered(f) ? (size_t)copies * v.node_count * 4 : 0;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t bytes = al(sz_desc) + al(sz_pref) + al(sz_cen) + al(sz_cb) + 2 * al(sz_keys) + 2 * al(sz_vals) +
                         al(sz_old) + al(sz_tasks) + al(sz_lv) + al(sz_order) + al(sz_q) + al(sort_bytes) +
                         al((size_t)n_ranges * 8);
    if (bytes > f->rebuild_bytes) {
        HIP_OK(f, hipStreamSynchronize(st));
        dfree(f->d_rebuild);
        f->rebuild_bytes = 0;
        if (hipMalloc(&f->d_rebuild, bytes) != hipSuccess) {
            f->d_rebuild = nullptr;
            return fail(f, SBGI_E_NOMEM, "rebuild scratch");
        }
        f->rebuild_bytes = bytes;
    }
    char *p = static_cast<char *>(f->d_rebuild);
    auto take = [&](size_t sz) {
        char *q = p;
        p += al(sz);
        return q;
    };
    uint4 *d_desc = reinterpret_cast<uint4 *>(take(sz_desc));
    uint32_t *d_pref = reinterpret_cast<uint32_t *>(take(sz_pref));
    float4 *d_cen = reinterpret_cast<float4 *>(take(sz_cen));
    uint32_t *d_cb = reinterpret_cast<uint32_t *>(take(sz_cb));
    unsigned long long *d_keys = reinterpret_cast<unsigned long long *>(take(sz_keys));
    unsigned long long *d_keys_s = reinterpret_cast<unsigned long long *>(take(sz_keys));
    uint32_t *d_vals = reinterpret_cast<uint32_t *>(take(sz_vals));
    uint32_t *d_vals_s = reinterpret_cast<uint32_t *>(take(sz_vals));
    uint32_t *d_old = reinterpret_cast<uint32_t *>(take(sz_old));  // 3 index words + material per triangle
    uint4 *d_tasks = reinterpret_cast<uint4 *>(take(sz_tasks));
    uint32_t *d_lv = reinterpret_cast<uint32_t *>(take(sz_lv));
    uint32_t *d_order = reinterpret_cast<uint32_t *>(take(sz_order));
    uint32_t *d_q = reinterpret_cast<uint32_t *>(take(sz_q));
    void *d_sort = take(sort_bytes);
    uint2 *d_tr = reinterpret_cast<uint2 *>(take((size_t)n_ranges * 8));
    int32_t rc = SBGI_OK;
    if ((rc = record_kev(f, 0)) != SBGI_OK || (rc = record_kev(f, 1)) != SBGI_OK) return rc;
    std::vector<uint32_t> cb_init((size_t)n_ranges * 6);
    for (uint32_t r = 0; r < n_ranges; r++)
        for (int q = 0; q < 3; q++) {
            cb_init[r * 6 + q] = 0xFF800000u;      // order-mapped +inf
            cb_init[r * 6 + 3 + q] = 0x007FFFFFu;  // order-mapped -inf
        }
    std::vector<uint2> trh(n_ranges);
    uint32_t tmax = 0;
    for (uint32_t r = 0; r < n_ranges; r++) {
        trh[r] = make_uint2(desc[r].z, desc[r].z + desc[r].w);
        tmax = std::max(tmax, desc[r].w);
    }
    HIP_OK(f, hipMemcpyAsync(d_desc, desc.data(), sz_desc, hipMemcpyHostToDevice, st));
    HIP_OK(f, hipMemcpyAsync(d_pref, prefix.data(), sz_pref, hipMemcpyHostToDevice, st));
    HIP_OK(f, hipMemcpyAsync(d_cb, cb_init.data(), sz_cb, hipMemcpyHostToDevice, st));
    HIP_OK(f, hipMemcpyAsync(d_tr, trh.data(), (size_t)n_ranges * 8, hipMemcpyHostToDevice, st));
    // the ranges' triangle rows, concatenated (the permutation gathers from this copy)
    uint32_t *d_old_idx = d_old, *d_old_mat = d_old + (size_t)n * 3;
    for (uint32_t r = 0; r < n_ranges; r++) {
        HIP_OK(f, hipMemcpyAsync(d_old_idx + (size_t)prefix[r] * 3, f->d_tri_index + (size_t)desc[r].z * 3,
                                 (size_t)desc[r].w * 12, hipMemcpyDeviceToDevice, st));
        HIP_OK(f, hipMemcpyAsync(d_old_mat + prefix[r], f->d_tri_mat + desc[r].z, (size_t)desc[r].w * 4,
                                 hipMemcpyDeviceToDevice, st));
    }
    HIP_OK(f, launch_rebuild_prepare(d_desc, d_pref, n_ranges, n, f->d_tri_index, f->d_vdata, d_cen, d_cb, d_keys,
                                     d_vals, d_keys_s, d_vals_s, d_sort, sort_bytes, 30u + range_bits, st));
    HIP_OK(f, launch_permute_tris(d_desc, d_pref, n_ranges, n, d_vals_s, d_old_idx, d_old_mat, f->d_tri_index,
                                  f->d_tri_mat, st));
    // top-down structure, one launch per depth over the device depth table: the roots are
    // depth 0's tasks, every later count is appended on the device
    std::vector<uint4> roots(n_ranges);
    for (uint32_t r = 0; r < n_ranges; r++) roots[r] = make_uint4(desc[r].x, r, 0u, (desc[r].y + 1) / 2);
    std::vector<uint32_t> lv_init((size_t)depths * 2 + 3, 0u);
    lv_init[1] = n_ranges;
    HIP_OK(f, hipMemcpyAsync(d_tasks, roots.data(), (size_t)n_ranges * 16, hipMemcpyHostToDevice, st));
    HIP_OK(f, hipMemcpyAsync(d_lv, lv_init.data(), sz_lv, hipMemcpyHostToDevice, st));
    HIP_OK(f, launch_build_depths(d_desc, d_pref, d_keys_s, d_tasks, cap, d_lv, depths, blocks.data(), f->d_nodes,
                                  d_order, st));
    // bounds bottom-up (deepest depth first)
    HIP_OK(f, launch_refit_depths(f->d_nodes, f->d_tri_index, f->d_vdata, d_order, d_lv, depths, cap, blocks.data(),
                                  st));
    // walk copies: positions top down (ordered), then every node of every copy
    if (ordered(f)) {
        HIP_OK(f, launch_copy_roots(d_desc, n_ranges, copies, d_q, v.node_count, st));
        HIP_OK(f, launch_copy_positions(f->d_nodes, d_order, d_lv, depths, cap, blocks.data(), copies, d_q,
                                        v.node_count, st));
    }
    HIP_OK(f, launch_write_copies(f->d_nodes, d_order, cap, d_lv + (size_t)depths * 2, copies, d_q, v.node_count,
                                  v.node_count, f->d_walk_nodes, f->d_walk_perm, st));
    HIP_OK(f, launch_pack_tris_ranges(f->d_tri_index, f->d_vdata, f->d_tri_mat, f->d_mats, v.tri_count, d_tr, n_ranges,
                                      tmax, f->d_tri, st));
    if ((rc = record_kev(f, 2)) != SBGI_OK) return rc;
    f->kev_blas = true;
    if ((rc = record_kev(f, 3)) != SBGI_OK || (rc = device_tlas(f, true)) != SBGI_OK || (rc = record_kev(f, 4)) != SBGI_OK)
        return rc;
    f->kev_tlas = true;
    // the new structure into the host mirrors
    for (uint32_t r = 0; r < n_ranges; r++) {
        HIP_OK(f, hipMemcpyAsync(f->h_nodes.data() + (size_t)desc[r].x * 8, f->d_nodes + (size_t)desc[r].x * 8,
                                 (size_t)desc[r].y * 32, hipMemcpyDeviceToHost, st));
        HIP_OK(f, hipMemcpyAsync(f->h_tri_index.data() + (size_t)desc[r].z * 3, f->d_tri_index + (size_t)desc[r].z * 3,
                                 (size_t)desc[r].w * 12, hipMemcpyDeviceToHost, st));
        HIP_OK(f, hipMemcpyAsync(f->h_tri_mat.data() + desc[r].z, f->d_tri_mat + desc[r].z, (size_t)desc[r].w * 4,
                                 hipMemcpyDeviceToHost, st));
    }
    // the depth table (its end: {nodes built, tasks past the bound, overflow flag})
    std::vector<uint32_t> lvh((size_t)depths * 2 + 3);
    HIP_OK(f, hipMemcpyAsync(lvh.data(), d_lv, lvh.size() * 4, hipMemcpyDeviceToHost, st));
    HIP_OK(f, hipStreamSynchronize(st));
    if (lvh[(size_t)depths * 2] != cap || lvh[(size_t)depths * 2 + 1] || lvh[(size_t)depths * 2 + 2])
        return fail(f, SBGI_E_LAYOUT, "rebuild node count mismatch");
    // the rebuilt ranges' refit plan straight from the depth table, deepest depth first (a
    // node's children are one depth deeper, so each level finds its children finished)
    auto depth_plan = [&](RefitPlan &pl, const uint32_t *key, size_t key_words) -> int32_t {
        if (pl.cap < nodes_total) {
            dfree(pl.d_order);
            pl.cap = 0;
            if (dalloc(&pl.d_order, nodes_total * 4) != hipSuccess) return fail(f, SBGI_E_NOMEM, "refit plan allocation");
            pl.cap = nodes_total;
        }
        pl.off.assign(1, 0u);
        for (uint32_t d = depths; d-- > 0;) {
            const uint32_t cnt = lvh[(size_t)d * 2 + 1];
            if (cnt)
                HIP_OK(f, hipMemcpyAsync(pl.d_order + pl.off.back(), d_order + lvh[(size_t)d * 2], (size_t)cnt * 4,
                                         hipMemcpyDeviceToDevice, st));
            pl.off.push_back(pl.off.back() + cnt);
        }
        pl.h_order.clear();
        pl.key.assign(key, key + key_words);
        return SBGI_OK;
    };
    // cached refit plans and the differentiable mode's triangle map follow the structure
    if ((rc = depth_plan(f->call_plan, blas_ranges, (size_t)n_ranges * 2)) != SBGI_OK) return rc;
    if (f->skin_ranges_n) {
        std::vector<uint32_t> kr(f->skin_plan.key);
        std::vector<std::pair<uint32_t, uint32_t>> skin_set;
        for (size_t i = 0; i + 1 < kr.size(); i += 2) skin_set.emplace_back(kr[i], kr[i + 1]);
        std::sort(skin_set.begin(), skin_set.end());
        f->skin_plan.key.clear();
        if (skin_set == nr)
            rc = depth_plan(f->skin_plan, kr.data(), kr.size());
        else
            rc = build_plan(f, kr.data(), (uint32_t)kr.size() / 2, f->skin_plan);
        if (rc != SBGI_OK) return rc;
    }
    f->diff_ready = f->diff_valid = false;
    return SBGI_OK;
}

int32_t sbgi_edit_timing(sbgi_field *f, float *ms_skin, float *ms_blas_refit, float *ms_tlas) {
    if (!f) return SBGI_E_INVALID;
    HIP_OK(f, hipSetDevice(f->desc.device));
    float s = 0.f, r = 0.f, t = 0.f;
    if (f->kev_blas) {
        HIP_OK(f, hipEventSynchronize(f->kev[2]));
        HIP_OK(f, hipEventElapsedTime(&s, f->kev[0], f->kev[1]));
        HIP_OK(f, hipEventElapsedTime(&r, f->kev[1], f->kev[2]));
    }
    if (f->kev_tlas) {
        HIP_OK(f, hipEventSynchronize(f->kev[4]));
        HIP_OK(f, hipEventElapsedTime(&t, f->kev[3], f->kev[4]));
    }
    if (ms_skin) *ms_skin = s;
    if (ms_blas_refit) *ms_blas_refit = r;
    if (ms_tlas) *ms_tlas = t;
    return SBGI_OK;
}

const char *sbgi_last_error(const sbgi_field *f) { return f ? f->err.c_str() : g_create_error.c_str(); }

int32_t sbgi_destroy(sbgi_field *f) {
    if (!f) return SBGI_OK;
    (void)hipSetDevice(f->desc.device);
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    drop_tick_graph(f);
    free_scene(f);
    for (auto &cs : f->casc) {
        dfree(cs.irr);
        dfree(cs.depth);
        dfree(cs.meta);
        dfree(cs.atlas);
        dfree(cs.datlas);
        dfree(cs.evis);
        dfree(cs.rough);
        dfree(cs.ratlas);
        dfree(cs.gnum);
        dfree(cs.gden);
        dfree(cs.gatlas);
    }
    dfree(f->rays);
    dfree(f->dirs);
    dfree(f->tw.hits);
    dfree(f->tw.vis);
    dfree(f->tw.queue.count);
    dfree(f->hit_ids);
    dfree(f->tables);
    dfree(f->counters);
    dfree(f->d_light_grid);
    dfree(f->rtables);
    for (auto &e : f->ring)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : f->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : f->tev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : f->dev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : f->kev)
        if (e) (void)hipEventDestroy(e);
    if (f->eev) (void)hipEventDestroy(f->eev);
    for (auto &e : f->sev)
        if (e) (void)hipEventDestroy(e);
    dfree(f->d_sample);
    for (int s = 0; s < sbgi_field::kStageSlots; s++) {
        if (f->stage_done[s]) (void)hipEventDestroy(f->stage_done[s]);
        if (f->stage_host[s]) (void)hipHostFree(f->stage_host[s]);
    }
    if (f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
    return SBGI_OK;
}

int32_t sbgi_classify(sbgi_field *f, uint32_t cascade) {
    if (!f) return SBGI_E_INVALID;
    if (cascade >= f->desc.cascade_count) return fail(f, SBGI_E_INVALID, "cascade out of range");
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    HIP_OK(f, hipSetDevice(f->desc.device));
    const CascadeState &cs = f->casc[cascade];
    const size_t need = (size_t)kClassifyRays * (1 + (size_t)cs.total);
    if (need > f->classify_cap) {
        HIP_OK(f, hipStreamSynchronize(f->stream));
        dfree(f->d_classify);
        f->classify_cap = 0;
        if (dalloc(&f->d_classify, need * sizeof(float4)) != hipSuccess)
            return fail(f, SBGI_E_NOMEM, "classification scratch");
        f->classify_cap = need;
    }
    // relocClamp = 0.45 x minCell, the smallest cell side (js/gi_probes.js:3153, 3364)
    const double gx = cs.d.grid_size[0], gy = cs.d.grid_size[1], gz = cs.d.grid_size[2];
    const double cm = std::max(1e-4, std::min({gx / std::max(1u, cs.d.res[0] - 1), gy / std::max(1u, cs.d.res[1] - 1),
                                               gz / std::max(1u, cs.d.res[2] - 1)}));
    const DevCascade c = make_cascade(f, cs, 0, cs.total);
    HIP_OK(f, launch_classify(make_scene(f), c, (float)(0.45 * cm), f->d_classify, f->d_classify + kClassifyRays,
                              cs.meta, f->tw.queue.count + trace_heads(2), f->stream));
    return SBGI_OK;
}

int32_t sbgi_query_rays(sbgi_field *f, const float *rays8, uint32_t n, float *out4) {
    if (!f || (n && (!rays8 || !out4))) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    if (n == 0) return SBGI_OK;
    HIP_OK(f, hipSetDevice(f->desc.device));
    float4 *din = nullptr, *dout = nullptr;
    if (dalloc(&din, (size_t)n * 32) != hipSuccess || dalloc(&dout, (size_t)n * 16) != hipSuccess) {
        dfree(din);
        return fail(f, SBGI_E_NOMEM, "query allocation failed");
    }
    hipError_t e = hipMemcpyAsync(din, rays8, (size_t)n * 32, hipMemcpyHostToDevice, f->stream);
    if (e == hipSuccess) e = launch_query(make_scene(f), din, dout, n, f->tw.queue.count + trace_heads(2), f->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out4, dout, (size_t)n * 16, hipMemcpyDeviceToHost, f->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
    dfree(din);
    dfree(dout);
    if (e != hipSuccess) return hip_fail(f, e, "sbgi_query_rays");
    return SBGI_OK;
}

int32_t sbgi_query_rays_noalloc(sbgi_field *f, const float *rays8, uint32_t n, float *out4) {
    if (!f || (n && (!rays8 || !out4))) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    if (n > SBGI_NOALLOC_MAX_RAYS)
        return fail(f, SBGI_E_CAPACITY, "sbgi_query_rays_noalloc takes at most SBGI_NOALLOC_MAX_RAYS (16384) rays per call");
    if (n == 0) return SBGI_OK;
    HIP_OK(f, hipSetDevice(f->desc.device));
    unsigned char *base = scratch_of(f->desc.device);
    if (!base) return fail(f, SBGI_E_STATE, "query scratch not set up on this device");
    uint32_t *heads = reinterpret_cast<uint32_t *>(base);
    float4 *din = reinterpret_cast<float4 *>(base + kNoAllocHeadBytes);
    float4 *dout = din + 2 * (size_t)SBGI_NOALLOC_MAX_RAYS;
    // the same kernels as sbgi_query_rays (launch_query clears the fetch counters first);
    // only where the staging and counters live differs
    hipError_t e = hipMemcpyAsync(din, rays8, (size_t)n * 32, hipMemcpyHostToDevice, f->stream);
    if (e == hipSuccess) e = launch_query(make_scene(f), din, dout, n, heads, f->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out4, dout, (size_t)n * 16, hipMemcpyDeviceToHost, f->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
    if (e != hipSuccess) return hip_fail(f, e, "sbgi_query_rays_noalloc");
    return SBGI_OK;
}

void sbgi_sample_params_default(sbgi_sample_params *p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->normal_bias = 0.04f;  // normalBias (js/gi_probes.js:314, 1043) x sampleBiasScale 1
    p->border_band = 0.15f;  // borderBand (js/gi_probes.js:350)
}

int32_t sbgi_sample_irradiance(sbgi_field *f, const sbgi_sample_params *p, const float *pos, const float *nvis,
                               const float *ndir, uint32_t n, float *out, float *ms) {
    if (!f || !p || (n && (!pos || !nvis || !out))) return SBGI_E_INVALID;
    const uint32_t nc = p->cascades ? p->cascades : f->desc.cascade_count;
    if (nc > f->desc.cascade_count) return fail(f, SBGI_E_INVALID, "more cascades than the field has");
    if (ms) *ms = 0.f;
    if (n == 0) return SBGI_OK;
    if ((uint64_t)n * 3 >= (1ull << 32)) return fail(f, SBGI_E_CAPACITY, "too many points");
    HIP_OK(f, hipSetDevice(f->desc.device));
    const size_t vec = (size_t)n * 12;
    const size_t need = vec * (ndir ? 4 : 3);
    if (need > f->sample_cap) {
        dfree(f->d_sample);
        f->sample_cap = 0;
        if (dalloc(&f->d_sample, need) != hipSuccess) return fail(f, SBGI_E_NOMEM, "sample buffer allocation failed");
        f->sample_cap = need;
    }
    float *d_pos = f->d_sample, *d_nvis = d_pos + 3 * (size_t)n, *d_out = d_nvis + 3 * (size_t)n;
    float *d_ndir = ndir ? d_out + 3 * (size_t)n : nullptr;
    SampleCascadeIn sc[SBGI_MAX_CASCADES];
    for (uint32_t c = 0; c < nc; c++) {
        const CascadeState &cs = f->casc[c];
        sc[c] = SampleCascadeIn{make_cascade(f, cs, 0, cs.total), cs.atlas, cs.datlas, cs.meta};
    }
    hipStream_t st = f->stream;
    HIP_OK(f, hipMemcpyAsync(d_pos, pos, vec, hipMemcpyHostToDevice, st));
    HIP_OK(f, hipMemcpyAsync(d_nvis, nvis, vec, hipMemcpyHostToDevice, st));
    if (ndir) HIP_OK(f, hipMemcpyAsync(d_ndir, ndir, vec, hipMemcpyHostToDevice, st));
    HIP_OK(f, hipEventRecord(f->sev[0], st));
    HIP_OK(f, launch_sample(sc, nc, *p, d_pos, d_nvis, d_ndir, n, d_out, st));
    HIP_OK(f, hipEventRecord(f->sev[1], st));
    HIP_OK(f, hipMemcpyAsync(out, d_out, vec, hipMemcpyDeviceToHost, st));
    HIP_OK(f, hipStreamSynchronize(st));
    if (ms) HIP_OK(f, hipEventElapsedTime(ms, f->sev[0], f->sev[1]));
    return SBGI_OK;
}

// Counting mode: the next ticks accumulate traversal counters (node visits, triangle
// tests, instance entries, shadow rays) for the algorithmic-bytes accounting.
int32_t sbgi_set_counting(sbgi_field *f, int32_t on) {
    if (!f) return SBGI_E_INVALID;
    f->count_mode = on != 0;
    HIP_OK(f, hipMemsetAsync(f->counters, 0, sizeof(Counters), f->stream));
    HIP_OK(f, hipStreamSynchronize(f->stream));
    return SBGI_OK;
}

// Stage timing: captured ticks record their stage events (the tick key holds the flag,
// so a change recaptures).
int32_t sbgi_set_stage_timing(sbgi_field *f, int32_t on) {
    if (!f) return SBGI_E_INVALID;
    f->stage_timing = on != 0;
    return SBGI_OK;
}


}  // extern "C"

// ---- differentiable mode ---------------------------------------------------------

namespace {

// Per TLAS slot: (offset into the albedo array, first pool triangle of its BLAS); entries
// are slot-major, then pool-triangle order within the BLAS.
int32_t diff_prepare(sbgi_field *f) {
    if (f->diff_ready) return SBGI_OK;
    const sbgi_scene_view &v = f->view;
    f->h_albedo_map.assign((size_t)v.inst_count * 2, 0);
    uint64_t off = 0;
    for (uint32_t i = 0; i < v.inst_count; i++) {
        const float *r = f->h_mats.data() + v.inst_base + (size_t)i * 28;
        const uint32_t root = (uint32_t)r[12], end = (uint32_t)r[13];
        uint32_t first = 0xFFFFFFFFu;
        uint64_t count = 0;
        for (uint32_t nd = root; nd < end; nd++) {
            const uint32_t pay = f->h_nodes[(size_t)nd * 8 + 7];
            if (pay == 0xFFFFFFFFu) continue;
            first = std::min(first, pay & 0x00FFFFFFu);
            count += pay >> 24;
        }
        if (first == 0xFFFFFFFFu) first = 0;
        f->h_albedo_map[2 * i] = (int32_t)off;
        f->h_albedo_map[2 * i + 1] = (int32_t)first;
        off += count;
    }
    if (off >= (1ull << 28)) return fail(f, SBGI_E_CAPACITY, "too many albedo entries");
    f->diff_albedo_n = (uint32_t)off;
    const CascadeState &cs = f->casc[0];
    const uint32_t n = cs.total * f->desc.rays_per_probe;
    const uint32_t cells = (std::max(cs.d.res[0], 2u) - 1u) * (std::max(cs.d.res[1], 2u) - 1u) *
                           (std::max(cs.d.res[2], 2u) - 1u);
    const uint32_t L = std::max(v.light_count, 1u);
    if (dalloc(&f->d_albedo, (size_t)std::max(f->diff_albedo_n, 1u) * 12) != hipSuccess ||
        dalloc(&f->d_galb, (size_t)std::max(f->diff_albedo_n, 1u) * 12) != hipSuccess ||
        dalloc(&f->d_light_scale, (size_t)L * 4) != hipSuccess || dalloc(&f->d_glight, (size_t)L * 4) != hipSuccess ||
        dalloc(&f->d_albedo_map, (size_t)std::max(v.inst_count, 1u) * 8) != hipSuccess ||
        dalloc(&f->d_atlas1, (size_t)cs.aw * cs.ah * 16) != hipSuccess ||
        dalloc(&f->d_gE, (size_t)cs.total * 108 * 4 * 2) != hipSuccess ||
        dalloc(&f->d_adj_scratch, adjoint_scratch_bytes(n, cells, v.light_count)) != hipSuccess)
        return fail(f, SBGI_E_NOMEM, "differentiable-mode allocation");
    HIP_OK(f, hipMemcpyAsync(f->d_albedo_map, f->h_albedo_map.data(), f->h_albedo_map.size() * 4,
                             hipMemcpyHostToDevice, f->stream));
    f->diff_ready = true;
    return SBGI_OK;
}

DevScene diff_scene(const sbgi_field *f) {
    DevScene s = make_scene(f);
    s.albedo = f->d_albedo;
    s.albedo_map = f->d_albedo_map;
    s.light_scale = f->d_light_scale;
    return s;
}

}  // namespace

extern "C" {

int32_t sbgi_diff_param_count(sbgi_field *f, uint32_t *albedo_entries, uint32_t *lights) {
    if (!f || !albedo_entries || !lights) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    HIP_OK(f, hipSetDevice(f->desc.device));
    int32_t rc = diff_prepare(f);
    if (rc != SBGI_OK) return rc;
    *albedo_entries = f->diff_albedo_n;
    *lights = f->view.light_count;
    return SBGI_OK;
}

int32_t sbgi_diff_albedo_map(sbgi_field *f, int32_t *map2, uint32_t n_slots) {
    if (!f || !map2) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    int32_t rc = diff_prepare(f);
    if (rc != SBGI_OK) return rc;
    if (n_slots != f->view.inst_count) return fail(f, SBGI_E_INVALID, "map size must be the instance count");
    memcpy(map2, f->h_albedo_map.data(), f->h_albedo_map.size() * 4);
    return SBGI_OK;
}

int32_t sbgi_diff_forward(sbgi_field *f, const sbgi_tick_params *p, const float *albedo, const float *light_scale,
                          float *irr_out) {
    if (!f || !p || !albedo || !irr_out) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    if (f->desc.depth_oct_res != 6) return fail(f, SBGI_E_INVALID, "the differentiable mode uses the 8x8 depth tile");
    if (f->has_scene && f->view.emitter_count)
        return fail(f, SBGI_E_INVALID, "the differentiable mode takes no emitter lights");
    HIP_OK(f, hipSetDevice(f->desc.device));
    int32_t rc = diff_prepare(f);
    if (rc != SBGI_OK) return rc;
    CascadeState &cs = f->casc[0];
    hipStream_t st = f->stream;
    sbgi_tick_params q = *p;
    q.cascade = 0;
    q.probe_offset = 0;
    q.updated_count = 0;
    q.needs_clear = 0;
    q.classify_strength = 0.f;
    q.filter_strength = 0.f;
    const DevCascade c = make_cascade(f, cs, 0, cs.total);
    const DevScene s = diff_scene(f);
    const uint32_t L = f->view.light_count;
    std::vector<float> ones(std::max(L, 1u), 1.f);
    HIP_OK(f, hipMemcpyAsync(f->d_albedo, albedo, (size_t)f->diff_albedo_n * 12, hipMemcpyHostToDevice, st));
    HIP_OK(f, hipMemcpyAsync(f->d_light_scale, light_scale ? light_scale : ones.data(), (size_t)std::max(L, 1u) * 4,
                             hipMemcpyHostToDevice, st));
    HIP_OK(f, hipEventRecord(f->dev[0], st));
    HIP_OK(f, launch_raydirs(f->dirs, c.rays, q.frame_jitter, st));
    HIP_OK(f, launch_weights(c, q, f->dirs, f->bt, SBGI_BLEND_FP32, st));
    const size_t irr_bytes = (size_t)cs.total * 64 * 16, dep_bytes = (size_t)cs.total * 64 * 8;
    // update 1: cleared state, zero bounce source; its upload writes the FP32 atlas A1
    HIP_OK(f, hipMemsetAsync(cs.irr, 0, irr_bytes, st));
    HIP_OK(f, hipMemsetAsync(cs.depth, 0, dep_bytes, st));
    HIP_OK(f, hipMemsetAsync(f->d_atlas1, 0, (size_t)cs.aw * cs.ah * 16, st));
    HIP_OK(f, launch_trace(s, c, q, f->dirs, cs.meta, cs.atlas, f->d_atlas1, f->tw, f->rays, nullptr, nullptr, nullptr,
                           st));
    HIP_OK(f, launch_blend(c, q, f->rays, f->bt, SBGI_BLEND_FP32, cs.irr, cs.depth, EmitterInject{}, st));
    HIP_OK(f, launch_upload(c, q, cs.irr, cs.depth, cs.atlas, cs.datlas, f->d_atlas1, st));
    // update 2: cleared state again, bouncing off A1
    HIP_OK(f, hipMemsetAsync(cs.irr, 0, irr_bytes, st));
    HIP_OK(f, hipMemsetAsync(cs.depth, 0, dep_bytes, st));
    HIP_OK(f, launch_trace(s, c, q, f->dirs, cs.meta, cs.atlas, f->d_atlas1, f->tw, f->rays, nullptr, nullptr, nullptr,
                           st));
    HIP_OK(f, launch_blend(c, q, f->rays, f->bt, SBGI_BLEND_FP32, cs.irr, cs.depth, EmitterInject{}, st));
    HIP_OK(f, hipEventRecord(f->dev[1], st));
    std::vector<float> irr((size_t)cs.total * 64 * 4);
    HIP_OK(f, hipMemcpyAsync(irr.data(), cs.irr, irr_bytes, hipMemcpyDeviceToHost, st));
    HIP_OK(f, hipStreamSynchronize(st));
    for (uint32_t pr = 0; pr < cs.total; pr++)
        for (uint32_t t = 0; t < 36; t++)
            for (int a = 0; a < 3; a++)
                irr_out[((size_t)pr * 36 + t) * 3 + a] = irr[((size_t)pr * 64 + (t / 6 + 1) * 8 + (t % 6 + 1)) * 4 + a];
    HIP_OK(f, hipEventElapsedTime(&f->ms_diff_fwd, f->dev[0], f->dev[1]));
    f->diff_p = q;
    f->diff_valid = true;
    return SBGI_OK;
}

int32_t sbgi_diff_adjoint(sbgi_field *f, const float *grad_irr, float *grad_albedo, float *grad_light_scale) {
    if (!f || !grad_irr || !grad_albedo) return SBGI_E_INVALID;
    if (!f->diff_valid) return fail(f, SBGI_E_STATE, "no differentiable forward to take the adjoint of");
    HIP_OK(f, hipSetDevice(f->desc.device));
    CascadeState &cs = f->casc[0];
    hipStream_t st = f->stream;
    const DevCascade c = make_cascade(f, cs, 0, cs.total);
    const DevScene s = diff_scene(f);
    const uint32_t L = f->view.light_count;
    float *gE2 = f->d_gE, *gE1 = f->d_gE + (size_t)cs.total * 108;
    HIP_OK(f, hipMemcpyAsync(gE2, grad_irr, (size_t)cs.total * 108 * 4, hipMemcpyHostToDevice, st));
    HIP_OK(f, hipEventRecord(f->dev[2], st));
    HIP_OK(f, hipMemsetAsync(f->d_galb, 0, (size_t)f->diff_albedo_n * 12, st));
    HIP_OK(f, hipMemsetAsync(f->d_glight, 0, (size_t)std::max(L, 1u) * 4, st));
    HIP_OK(f, launch_adjoint(s, c, f->diff_p, f->dirs, f->bt, f->tw, f->d_atlas1, gE2, gE1, f->d_galb, f->d_glight,
                             f->d_adj_scratch, st));
    HIP_OK(f, hipEventRecord(f->dev[3], st));
    HIP_OK(f, hipMemcpyAsync(grad_albedo, f->d_galb, (size_t)f->diff_albedo_n * 12, hipMemcpyDeviceToHost, st));
    if (grad_light_scale && L)
        HIP_OK(f, hipMemcpyAsync(grad_light_scale, f->d_glight, (size_t)L * 4, hipMemcpyDeviceToHost, st));
    HIP_OK(f, hipStreamSynchronize(st));
    HIP_OK(f, hipEventElapsedTime(&f->ms_diff_adj, f->dev[2], f->dev[3]));
    return SBGI_OK;
}

int32_t sbgi_diff_timing(sbgi_field *f, float *ms_forward, float *ms_adjoint) {
    if (!f || !ms_forward || !ms_adjoint) return SBGI_E_INVALID;
    *ms_forward = f->ms_diff_fwd;
    *ms_adjoint = f->ms_diff_adj;
    return SBGI_OK;
}


// ---- walk schedule and autotuner --------------------------------------------------

int32_t sbgi_get_walk_tuning(sbgi_field *f, sbgi_walk_tuning *out) {
    if (!f || !out) return SBGI_E_INVALID;
    *out = f->tuning;
    return SBGI_OK;
}

int32_t sbgi_set_walk_tuning(sbgi_field *f, const sbgi_walk_tuning *t) {
    if (!f || !t) return SBGI_E_INVALID;
    if (t->tri_ratio < 1 || t->tri_ratio > 64 || t->inst_ratio < 1 || t->inst_ratio > 64 || t->refill_min < 1 ||
        t->refill_min > 64 || t->fetch_chunk < 1 || t->fetch_chunk > 1024 || t->shadow_lists > 1 ||
        (t->shadow_copies != 1 && t->shadow_copies != 8))
        return fail(f, SBGI_E_INVALID, "walk tuning out of range");
    HIP_OK(f, hipSetDevice(f->desc.device));
    f->tuning = *t;
    if (f->tuning.shadow_lists && f->sl_stale) return build_shadow_lists(f);
    return SBGI_OK;
}

int32_t sbgi_autotune(sbgi_field *f, const sbgi_tick_params *p, sbgi_walk_tuning *best, float *best_ms) {
    if (!f || !p) return SBGI_E_INVALID;
    if (!f->has_scene) return fail(f, SBGI_E_STATE, "no scene");
    if (p->cascade >= f->desc.cascade_count) return fail(f, SBGI_E_INVALID, "cascade index out of range");
    HIP_OK(f, hipSetDevice(f->desc.device));
    CascadeState &cs = f->casc[p->cascade];
    const uint32_t updated = p->updated_count ? std::min(p->updated_count, cs.total) : cs.total;
    const DevCascade c = make_cascade(f, cs, p->probe_offset % cs.total, updated);
    hipStream_t st = f->stream;
    HIP_OK(f, launch_raydirs(f->dirs, c.rays, p->frame_jitter, st));
    f->diff_valid = false;
    auto time_trace = [&](const sbgi_walk_tuning &t, float &ms) -> int32_t {
        const sbgi_walk_tuning keep = f->tuning;
        f->tuning = t;
        if (t.shadow_lists && f->sl_stale) {
            const int32_t rb = build_shadow_lists(f);
            if (rb != SBGI_OK) return rb;
        }
        const DevScene s = make_scene(f);
        f->tuning = keep;
        float best_run = 1e30f;
        for (int rep = 0; rep < 2; rep++) {
            HIP_OK(f, hipEventRecord(f->dev[0], st));
            HIP_OK(f, launch_trace(s, c, *p, f->dirs, cs.meta, cs.atlas, nullptr, f->tw, f->rays, nullptr, nullptr,
                                   nullptr, st));
            HIP_OK(f, hipEventRecord(f->dev[1], st));
            HIP_OK(f, hipEventSynchronize(f->dev[1]));
            float m = 0.f;
            HIP_OK(f, hipEventElapsedTime(&m, f->dev[0], f->dev[1]));
            best_run = std::min(best_run, m);
        }
        ms = best_run;
        return SBGI_OK;
    };
    sbgi_walk_tuning cur = f->tuning;
    float cur_ms = 0.f;
    int32_t rc = time_trace(cur, cur_ms);
    if (rc != SBGI_OK) return rc;
    static const uint32_t ratios[] = {2, 4, 8, 16}, refills[] = {4, 8, 16, 32}, chunks[] = {32, 64, 128, 256};
    static const uint32_t onoff[] = {0, 1}, copies[] = {1, 8};
    for (int pass = 0; pass < 2; pass++) {
        for (int param = 0; param < 6; param++) {
            // the shadow lists are a candidate only where the field has them, one shadow copy
            // only where there are eight
            if ((param == 4 && !f->sl.list) || (param == 5 && !ordered(f))) continue;
            const uint32_t *cand = param < 2    ? ratios
                                   : param == 2 ? refills
                                   : param == 3 ? chunks
                                   : param == 4 ? onoff
                                                : copies;
            for (int i = 0; i < (param >= 4 ? 2 : 4); i++) {
                sbgi_walk_tuning t = cur;
                uint32_t *slot = param == 0   ? &t.tri_ratio
                                 : param == 1 ? &t.inst_ratio
                                 : param == 2 ? &t.refill_min
                                 : param == 3 ? &t.fetch_chunk
                                 : param == 4 ? &t.shadow_lists
                                              : &t.shadow_copies;
                if (*slot == cand[i]) continue;
                *slot = cand[i];
                float ms = 0.f;
                if ((rc = time_trace(t, ms)) != SBGI_OK) return rc;
                if (ms < cur_ms * 0.99f) {
                    cur = t;
                    cur_ms = ms;
                }
            }
        }
    }
    f->tuning = cur;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return SBGI_OK;
}

}  // extern "C"
