// traverse2.h — wave-scheduled two-level traversal: one ray per lane as a small state
// machine, advanced by phases that the whole wave agrees on.
//
// The walk is the reference's (traverseClosest/blasClosest and traverseAny/blasAny,
// js/spectral_traverse.js:354-569): threaded pre-order nodes, i+1 on a box hit, the miss
// link otherwise, Moller-Trumbore with strict '<' on t, a TLAS leaf's instances walked to
// completion in slot order before the next TLAS node. The box test is the reference's
// slab formula with a correctly rounded 1/d, so every box, triangle and instance decision
// is the one the per-ray walk makes; only the schedule differs.
//
// Lane states:
//   NODE  test the box of node `cur` of the current level (TLAS in world space or the
//         instance's BLAS in local space) and move to i+1 / the miss link; a hit leaf
//         moves the lane to TRI (BLAS leaf) or INST (TLAS leaf).
//   TRI   one triangle of the pending leaf range per step.
//   INST  enter the next active instance of the pending TLAS leaf (local ray, inverse
//         direction, BLAS range), or return to the TLAS when the leaf is exhausted.
//   DONE  the ray has its answer and waits to be retired; IDLE: no ray.
// A lane is in NODE only while cur < end: the step that moves the cursor past the end of
// its level takes the level exit itself (BLAS: INST, TLAS: DONE), so no step is spent on
// an exhausted level. The cursor only moves forward (a threaded layout's i+1 and miss links
// are larger than i, and the step clamps them so), which bounds every walk.
// A wave runs the NODE phase whenever a lane needs it and the TRI / INST phases only
// when enough lanes are waiting in them (or nothing else is left), so the expensive
// phases run with most lanes busy instead of once per straggler. Finished lanes are
// refilled from the ray source, so a wave never idles on one long ray.
#pragma once
#include "traverse.h"

enum : uint32_t { ST_NODE = 0u, ST_TRI = 1u, ST_INST = 2u, ST_DONE = 3u, ST_IDLE = 4u };

struct LaneRay {
    float3 wo, wd, wid;   // world ray and its inverse direction
    float3 o, d, id;      // ray of the current level (world or instance-local)
    float3 noid;          // -o * id of the current level (the any-hit walk's fused slab test)
    float tmax;           // closest: best t so far; any: the shadow ray's extent
    uint32_t cur, end;    // node cursor and end of the current level
    uint32_t lend;        // end of the top level: TLAS nodes, or the ray's candidate list
    uint2 lquad;          // candidate lists: the 64-bit word of list entries holding entry
                          // `cur` (entries 4j .. 4j + 3, lowest first)
    uint32_t tres;        // TLAS node to resume at after the pending instances
    // pending ranges as leaf payload words, (count << 24) | offset: the instance slots of
    // the current TLAS leaf and the triangles of the current BLAS leaf still to take; a
    // step takes the offset and moves the word to (count - 1, offset + 1)
    uint32_t ipend, tpend;
    uint32_t base;        // node offset of the BLAS copy being walked (octant copy * nodes)
    uint32_t inst;        // instance being walked
    uint32_t level;       // 0 TLAS, 1 BLAS
    uint32_t state;
    float dsign;
    int hit_tri, hit_inst;
    uint32_t tag;         // caller's payload (shadow rays: (ray << 6) | light)
    bool any;             // any-hit walk (runtime mode of the query kernel)
    bool blocked;
};

// Inverse direction of the slab test: correctly rounded (the closest-hit walk, so box
// decisions are the FP32 restatement's), or the hardware reciprocal (RCP, 1 ulp: the
// any-hit walk of shadow rays, where a box decision can then differ only for a ray that
// grazes a box face, and a visibility answer only if it also grazes an occluder's edge;
// the reference's own WGSL division is specified to 2.5 ulp).
template <bool RCP = false>
__device__ __forceinline__ float3 ray_inv(float3 d) {
    if (RCP) return f3(__builtin_amdgcn_rcpf(d.x), __builtin_amdgcn_rcpf(d.y), __builtin_amdgcn_rcpf(d.z));
    return f3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
}

// Start a walk at the TLAS root (a scene has at least one TLAS node: the upload checks it).
template <bool RCP = false>
__device__ __forceinline__ void lane_begin(const DevScene &s, LaneRay &r, float3 ro, float3 rd, float tmax) {
    r.wo = ro;
    r.wd = rd;
    r.wid = ray_inv<RCP>(rd);
    r.o = ro;
    r.d = rd;
    r.id = r.wid;
    r.noid = f3(-(ro.x * r.wid.x), -(ro.y * r.wid.y), -(ro.z * r.wid.z));
    r.tmax = tmax;
    r.cur = 0u;
    r.end = s.tlas_count;
    r.lend = s.tlas_count;
    r.tres = 0u;
    r.ipend = r.tpend = 0u;
    r.base = s.tlas_oct ? octant_of(rd, 8u) * s.tlas_count : 0u;
    r.inst = 0u;
    r.level = 0u;
    r.state = ST_NODE;
    r.dsign = 1.f;
    r.hit_tri = -1;
    r.hit_inst = -1;
    r.blocked = false;
}

// Walk kinds: closest hit, any hit, or per ray (LaneRay::any, used by the query kernel).
enum : int { WALK_CLOSEST = 0, WALK_ANY = 1, WALK_MIXED = 2 };

// The step functions compute every new field value first and then assign each field
// exactly once, unconditionally: stores to different fields on different branches can be
// merged by the compiler into one store through a pointer phi, which forces the lane
// state out of registers into scratch memory.

// NODE: one box test at the current level (cur < end). Both levels use the 32-byte node
// {min.xyz, miss}, {max.xyz, payload} with payload 0xFFFFFFFF for interior nodes, else
// (count << 24) | offset (triangles for a BLAS leaf, instance slots for a TLAS leaf).
// BLAS nodes are addressed by a 32-bit byte offset from the pool: the upload limits a
// copy to fewer than 2^24 nodes, so 8 copies stay below 4 GiB.
// Exit-distance scale of the watertight walks' box tests: 1 + 4 gamma_3 (= 1 + 6 ulps,
// gamma_n = n u / (1 - n u)), twice the margin that covers the roundings of (b - o) * (1/d)
// (Ize 2013, "Robust BVH ray traversal", JCGT 2(2)), so a ray meeting a triangle on the
// face of its box is never culled.
constexpr float kRobustFar = 1.0000007152557373f;

// LIST: the top level is the ray's candidate list (k_shadow.hip): entry `cur` names an
// instance slot, tested as a one-instance leaf whose miss link is the next entry.
template <bool COUNT, bool LDS, bool FMA, bool WT = false, bool TOCT = false, bool LIST = false>
__device__ __forceinline__ void node_step(const DevScene &s, LaneRay &r, TravCounters<COUNT> &cnt) {
    uint4 a, b;
    if (LIST && !r.level) {
        // entries are read four per 64-bit word: the step of a word's last entry loads the
        // next word, so a list costs a quarter as many reads, each issued a step ahead
        const uint32_t q = r.cur & 3u, w = (q & 2u) ? r.lquad.y : r.lquad.x;
        const uint32_t slot = (q & 1u) ? (w >> 16) : (w & 0xFFFFu);
        if (q == 3u) r.lquad = reinterpret_cast<const uint2 *>(s.sl_list)[min(r.cur + 1u, r.lend - 1u) >> 2];
        list_box<LDS>(s, slot, a, b);
        a.w = r.cur + 1u;
        b.w = (1u << 24) | slot;
        cnt.add_tlas();
    } else if (r.level || (TOCT && s.tlas_oct)) {
        // BLAS copy, or (closest-hit walk) the near-first TLAS copy of the ray's octant
        const uint4 *pool = r.level ? s.nodes : s.tlas_oct;
        const uint4 *n = reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(pool) +
                                                         ((r.base + r.cur) << 5));
        a = n[0];
        b = n[1];
        if (r.level) cnt.add_blas();
        else cnt.add_tlas();
    } else {
        tlas_node<LDS>(s, r.cur, a, b);
        cnt.add_tlas();
    }
    const float3 bmin = f3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
    const float3 bmax = f3(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z));
    float tn, tf;
    if (FMA) slab_fma(bmin, bmax, r.noid, r.id, tn, tf);
    else slab(bmin, bmax, r.o, r.id, tn, tf);
    if (WT && s.watertight) tf = tf * kRobustFar;
    // tf >= max(tn, 0) as two compares: tn and tf are never NaN (a unit direction has a
    // finite plane distance on some axis, and fminf/fmaxf pass over NaN ones)
    const bool hit = tf >= tn && tf >= 0.f && tn < r.tmax;
    const bool descend = hit && b.w == 0xFFFFFFFFu;
    // an empty leaf (count 0) is passed over like a miss: TRI / INST never see an empty range
    const bool leaf = hit && !descend && (b.w >> 24) != 0u;
    // i+1 on a descent, else the miss link; masked rather than selected, so the compiler
    // cannot sink the miss link's load into a branch of its own (a second round trip)
    const uint32_t next = max(a.w & (descend ? 0u : ~0u), r.cur + 1u);
    // leaf: TRI (BLAS) or INST (TLAS); exhausted level: INST (BLAS) or DONE (TLAS)
    const uint32_t state = leaf ? ST_INST - r.level : (next < r.end ? ST_NODE : ST_DONE - r.level);
    const bool blas_leaf = leaf && r.level, tlas_leaf = leaf && !r.level;
    r.cur = next;
    r.tpend = blas_leaf ? b.w : r.tpend;
    r.ipend = tlas_leaf ? b.w : r.ipend;
    r.tres = tlas_leaf ? next : r.tres;
    r.state = state;
}

// TRI: one Moller-Trumbore test of the pending leaf range (or, in walks compiled with WT
// for a watertight scene, one watertight test of its stored vertices).
template <int MODE, bool COUNT, bool WT>
__device__ __forceinline__ void tri_step(const DevScene &s, LaneRay &r, TravCounters<COUNT> &cnt) {
    const bool ANY = MODE == WALK_MIXED ? r.any : MODE == WALK_ANY;
    const uint32_t tri = r.tpend & 0x00FFFFFFu;
    const float4 *tp = s.tri + (size_t)tri * 3;
    const float4 A = tp[0], B = tp[1], C = tp[2];
    cnt.add_tri();
    float th = 0.f, det = 0.f;
    bool hit;
    if (WT && s.watertight) {  // uniform over the launch
        const uint32_t *ti = s.tri_index + (size_t)tri * 3;
        const float4 *vd = reinterpret_cast<const float4 *>(s.vdata);  // 2 float4 per vertex row
        const float4 v0 = vd[(size_t)ti[0] * 2], v1 = vd[(size_t)ti[1] * 2], v2 = vd[(size_t)ti[2] * 2];
        hit = wt_test(xyz(v0), xyz(v1), xyz(v2), r.o, wt_setup(r.d), th, det);
    } else {
        hit = mt_test<MODE == WALK_ANY>(A, B, C, r.o, r.d, th, det);
    }
    const bool geo = hit && th > SBGI_RAY_EPS && th < r.tmax;
    const uint32_t flags = __float_as_uint(A.w);
    const bool accept = geo && side_ok(flags, det * r.dsign) && (flags & (ANY ? 8u : 4u));
    const bool block = ANY && accept, better = !ANY && accept;
    const uint32_t next = r.tpend + 1u - (1u << 24);  // offsets stay below 2^24
    r.tmax = better ? th : r.tmax;
    r.hit_tri = better ? (int)tri : r.hit_tri;
    r.hit_inst = better ? (int)r.inst : r.hit_inst;
    r.blocked = r.blocked || block;
    // leaf exhausted: on to the node after it, or back to INST when the BLAS is done
    r.state = block ? ST_DONE : ((next >> 24) ? ST_TRI : (r.cur < r.end ? ST_NODE : ST_INST));
    r.tpend = next;
}

// INST: enter the next instance slot of the pending TLAS leaf, or go back to the TLAS.
// Straight-line with selects: every INST lane reads a slot's rows (slot 0 when returning)
// and computes its local ray, so the compiler has no nested joins whose register copies
// would cost more than the selects. An inactive slot is entered with an empty range, which
// sends the lane straight back to INST for the next slot.
// Component-wise select: a select of whole float3 values can become a select of their
// addresses, which moves the lane state out of registers into scratch memory.
__device__ __forceinline__ float3 sel3(bool c, float3 a, float3 b) {
    return f3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z);
}

template <bool COUNT, bool LDS, bool RCP, bool TOCT = false, bool LIST = false>
__device__ __forceinline__ void inst_step(const DevScene &s, LaneRay &r, TravCounters<COUNT> &cnt) {
    const bool enter = (r.ipend >> 24) != 0u;
    const uint32_t slot = enter ? r.ipend & 0x00FFFFFFu : 0u;
    if (enter) cnt.add_inst();
    float4 r0, r1, r2, r3;
    inst_rows<LDS>(s, slot, r0, r1, r2, r3);
    const bool active = r3.w > 0.5f;
    const float3 lo = inst_point(r0, r1, r2, r.wo), ld = inst_dir(r0, r1, r2, r.wd);
    const float3 lid = ray_inv<RCP>(ld);
    const uint32_t c0 = active ? (uint32_t)r3.x : 0u, c1 = active ? (uint32_t)r3.y : 0u;
    r.o = sel3(enter, lo, r.wo);
    r.d = sel3(enter, ld, r.wd);
    r.id = sel3(enter, lid, r.wid);
    r.noid = f3(-(r.o.x * r.id.x), -(r.o.y * r.id.y), -(r.o.z * r.id.z));
    r.dsign = enter ? r3.z : r.dsign;
    const uint32_t top_end = LIST ? r.lend : s.tlas_count;
    r.cur = enter ? c0 : r.tres;
    r.end = enter ? c1 : top_end;
    // back on the TLAS: the near-first copy of the world ray's octant (closest-hit walk)
    const uint32_t tbase = TOCT && s.tlas_oct ? octant_of(r.wd, 8u) * s.tlas_count : r.base;
    r.base = enter ? octant_of(ld, s.blas_octants) * s.node_count : tbase;
    r.inst = enter ? slot : r.inst;
    r.level = enter ? 1u : 0u;
    // an empty range (or an inactive slot): on to the next slot
    r.state = enter ? (c0 < c1 ? ST_NODE : ST_INST) : (r.tres < top_end ? ST_NODE : ST_DONE);
    r.ipend = enter ? r.ipend + 1u - (1u << 24) : r.ipend;
}

// Wave schedule (DevScene::tri_ratio, inst_ratio, refill_min, fetch_chunk; the defaults
// came from a lockstep simulation of the instanced workload, and sbgi_autotune refines them
// per scene). A TRI or INST phase runs when the lanes waiting in it are at least 1/ratio of
// the lanes still walking nodes (so always once no lane is in NODE); retired lanes are
// refilled when at least refill_min are waiting, or nothing else is left.
//
// Ray fetching. One counter serialises every atomic on it (~10 ns each on MI355X), so the
// ray range is cut into kFetchHeads slices, each with its own counter kHeadStride words
// (128 bytes) apart, and a wave takes fetch_chunk rays per atomic. A wave starts on slice
// (wave id % kFetchHeads) and moves to the next slice when its slice is drained.

// Ray sources of walk_wave: load(scene, index, lane ray), store(index, lane ray), and
// kList (the top level is the ray's candidate list, which is never empty: a ray that can
// reach no instance gets a list of one box that no ray hits).
// walk_wave runs one wave until all n rays of the source are taken and walked; `heads`
// are the source's kFetchHeads fetch counters (zero before the launch), shared by every
// wave of the grid.
template <int MODE, bool COUNT, bool LDS, bool WT, class Src>
__device__ __forceinline__ void walk_wave(const DevScene &s, const Src &src, uint32_t n, uint32_t *heads,
                                          uint32_t wave_id, TravCounters<COUNT> &cnt) {
    constexpr bool LIST = Src::kList;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = lane ? (~0ull >> (64u - lane)) : 0ull;
    LaneRay r;
    r.state = ST_DONE;
    uint32_t idx = 0;
    bool have = false, drained = false;
    uint32_t h = wave_id % kFetchHeads, tried = 0, cnext = 0, cend = 0;
    // clamped so that a zero-initialised schedule still makes progress
    const int tri_ratio = (int)max(s.tri_ratio, 1u), inst_ratio = (int)max(s.inst_ratio, 1u);
    const int refill_min = (int)max(s.refill_min, 1u);
    const uint32_t chunk = max(s.fetch_chunk, 1u);
    for (;;) {
        const uint64_t done = __ballot(r.state == ST_DONE);
        uint64_t live = __ballot(r.state < ST_DONE);
        if (done && (__popcll(done) >= refill_min || !live)) {
            if (COUNT && lane == 0) cnt.refill_w++;
            if (r.state == ST_DONE) {
                if (have) src.store(idx, r);
                have = false;
            }
            const uint32_t want = (uint32_t)__popcll(done), rank = (uint32_t)__popcll(done & lt);
            const int leader = __ffsll((unsigned long long)done) - 1;
            uint32_t got = 0;
            bool take_ray = false;  // this lane is given ray idx in this refill
            while (got < want && !drained) {
                if (cnext >= cend) {
                    const uint32_t s0 = (uint32_t)(((uint64_t)n * h) / kFetchHeads);
                    const uint32_t s1 = (uint32_t)(((uint64_t)n * (h + 1u)) / kFetchHeads);
                    uint32_t off = 0;
                    if ((int)lane == leader) off = atomicAdd(heads + (size_t)h * kHeadStride, chunk);
                    off = (uint32_t)__shfl((int)off, leader);
                    if (off >= s1 - s0) {
                        h = h + 1u == kFetchHeads ? 0u : h + 1u;
                        drained = ++tried >= kFetchHeads;
                        continue;
                    }
                    cnext = s0 + off;
                    cend = min(cnext + chunk, s1);
                }
                const uint32_t take = min(want - got, cend - cnext);
                if (r.state == ST_DONE && rank >= got && rank < got + take) {
                    idx = cnext + (rank - got);
                    take_ray = true;
                }
                cnext += take;
                got += take;
            }
            // one load after the fetch loop, so none of its counters is live across it
            if (take_ray) {
                src.load(s, idx, r);
                have = true;
            }
            if (r.state == ST_DONE) r.state = ST_IDLE;
            live = __ballot(r.state < ST_DONE);
        }
        if (!live) break;
        if (__ballot(r.state == ST_NODE)) {
            if (COUNT && lane == 0) cnt.node_w++;
            if (r.state == ST_NODE) node_step<COUNT, LDS, MODE == WALK_ANY, WT, MODE == WALK_CLOSEST, LIST>(s, r, cnt);
        }
        uint64_t mn = __ballot(r.state == ST_NODE);
        const uint64_t mt = __ballot(r.state == ST_TRI);
        if (mt && __popcll(mt) * tri_ratio >= __popcll(mn)) {
            if (COUNT && lane == 0) cnt.tri_w++;
            if (r.state == ST_TRI) tri_step<MODE, COUNT, WT>(s, r, cnt);
        }
        mn = __ballot(r.state == ST_NODE);
        const uint64_t mi = __ballot(r.state == ST_INST);
        if (mi && __popcll(mi) * inst_ratio >= __popcll(mn)) {
            if (COUNT && lane == 0) cnt.inst_w++;
            if (r.state == ST_INST) inst_step<COUNT, LDS, MODE == WALK_ANY, MODE == WALK_CLOSEST, LIST>(s, r, cnt);
        }
    }
}
