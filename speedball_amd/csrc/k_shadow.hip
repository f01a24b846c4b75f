// k_shadow.hip — light-space candidate lists of the shadow-ray walk.
//
// NEE shadow rays (traverseAny, js/spectral_traverse.js:480-569, called per light from
// traceKernel, js/gi_probes.js:2129-2176) are ~80 % of the rays of an update, and in an
// open instanced scene a binary TLAS walk spends as many box tests finding the few
// instances such a ray passes (~37 TLAS nodes per ray on the 256-instance workload) as
// the BLAS walks inside them. Every shadow ray ends at its light, so it lies in the cone
// of directions from a point or spot light that contains its origin, or on the line
// through its origin along a directional light. Per light, this file lists for every cell
// of those directions (a 6 x 16 x 16 cube map around the light) or of the plane across the
// light direction (a 32 x 32 grid) the instance slots whose padded world box reaches the
// cell. The walk (traverse2.h, ShadowListSrc in k_trace.hip) takes its ray's list and tests
// each listed box as a one-instance TLAS leaf. The answer is the any-hit walk's: a list
// holds every instance the ray can reach, and which instances are tested first cannot
// change whether an occluder exists. A cell that no box reaches lists a NaN box, which no
// ray hits, so every list has an entry.
//
// The order does change how soon a blocked ray stops: each list holds its slots nearest the
// light first (a point or spot light: by the squared distance from the light to the box;
// directional: by the box centre's height along the direction toward the light; ties by
// slot). A box near the light lies between the light and most origins of its cell, so it
// is the likeliest occluder: on the instanced workload (256 instances, 16 point lights)
// this order cuts the shadow walk's BLAS node visits from 425 to 336 M and its time by a
// fifth against slot order (farthest first: 526 M).
//
// Conservative by construction:
// - a box's cells are widened by kFaceMargin in face coordinates (directional: kDirMargin
//   of the scene's largest coordinate), far more than the rounding of a ray's direction
//   and of the walk's cell arithmetic, except within kNearScale of the light's reach, where
//   a rounded direction can swing further: boxes that close to the light are listed in
//   every cell;
// - a box that straddles a face's plane through the light gets the unbounded side of that
//   face; one that holds the light gets every cell;
// - a ranged point or spot light skips boxes beyond its range (plus margins): its shadow
//   rays exist only within it (the gate, js/gi_probes.js:2129-2142).
//
// Build, all on the stream, no host sync (sbgi_field rebuilds it whenever instance boxes or
// light records change):
//   k_sl_proj   one block per light: projection record (cube-map centre, or the plane
//               basis and grid extent over every box corner) and the near radius;
//   k_sl_rects  per (light, slot): the box's cell rectangle on each face;
//   k_sl_count  per (light, cell): how many rectangles hold the cell, as an offset within
//               its block of cells, and the block's total;
//   k_sl_boxes  the list boxes: every slot's padded world box, then the NaN box;
//   k_sl_scan   one block: the blocks' exclusive offsets after the identity list [0, slots)
//               and the NaN box's entry, and the total;
//   k_sl_order  per (light, slot): the slot's rank in the light's order (a rank sort over
//               the light's slots, staged through LDS), written as order[light][rank];
//   k_sl_fill   per (light, cell): the slots in the light's order and the record
//               {first, end}. When the lists exceed the capacity every record becomes the
//               identity list (every slot): slower, still exact.
// The lists depend only on boxes and lights, are deterministic (a total order, no atomics),
// and cost ~2 x lights x cells x slots rectangle tests and lights x slots^2 comparisons to
// build.
#include "kernels.h"

namespace {

constexpr uint32_t kSlBlock = 256;
constexpr float kFaceMargin = 2e-3f;   // cube-map cells widened by this much of u, v in [-1, 1]
constexpr float kNearScale = 1e-2f;    // near radius: this fraction of the light's farthest box corner
constexpr float kDirMargin = 1e-4f;    // directional grid: this fraction of the largest |coordinate|
constexpr float kRangeSlack = 1e-3f;   // ranged lights: relative slack on the range

// rectangle of cells {u0, u1, v0, v1} packed in bytes; u0 > u1 is empty
__host__ __device__ constexpr uint32_t pack_rect(uint32_t u0, uint32_t u1, uint32_t v0, uint32_t v1) {
    return u0 | (u1 << 8) | (v0 << 16) | (v1 << 24);
}
constexpr uint32_t kEmptyRect = pack_rect(1, 0, 1, 0);
constexpr uint32_t kFullFace = pack_rect(0, kShadowFace - 1, 0, kShadowFace - 1);

__device__ __forceinline__ bool in_rect(uint32_t r, uint32_t u, uint32_t v) {
    return u >= (r & 0xFFu) && u <= ((r >> 8) & 0xFFu) && v >= ((r >> 16) & 0xFFu) && v <= (r >> 24);
}

__device__ __forceinline__ float3 box_lo(const float *b) { return f3(b[0], b[1], b[2]); }
__device__ __forceinline__ float3 box_hi(const float *b) { return f3(b[3], b[4], b[5]); }
__device__ __forceinline__ float comp3(float3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// light kind: point / spot (cube map), directional (plane grid), or none (type >= 2.5:
// emitters never queue shadow rays)
__device__ __forceinline__ float light_kind(const float *L) {
    return L[0] < 0.5f ? kShadowKindDir : (L[0] < 2.5f ? kShadowKindPoint : kShadowKindAll);
}

// the direction toward a directional light as light_vector takes it, and a basis across it
__device__ __forceinline__ void dir_basis(const float *L, float3 &e1, float3 &e2) {
    float3 w = f3(-L[4], -L[5], -L[6]);
    const float n = sqrtf(dot3(w, w));
    w = n > 0.f ? f3(w.x / n, w.y / n, w.z / n) : f3(0.f, 0.f, 1.f);
    const float3 h = fabsf(w.x) < 0.9f ? f3(1.f, 0.f, 0.f) : f3(0.f, 1.f, 0.f);
    e1 = normalize3(cross3(w, h));
    e2 = cross3(w, e1);
}

template <class Op>
__device__ __forceinline__ float block_reduce(float v, float *sh, Op op) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t s = kSlBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = op(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const float r = sh[0];
    __syncthreads();
    return r;
}

// proj[l] = {kind, near radius, range cut, 0}, then
//   point / spot: {position, 0};
//   directional:  {e1, u0}, {e2, v0}, {su, sv, margin, 0}
__global__ void __launch_bounds__(kSlBlock) k_sl_proj(const float *__restrict__ lights, const float *__restrict__ box,
                                                      uint32_t n_inst, float4 *__restrict__ proj) {
    __shared__ float sh[kSlBlock];
    const uint32_t l = blockIdx.x;
    const float *L = lights + (size_t)l * SBGI_LIGHT_STRIDE;
    const float kind = light_kind(L);
    float4 *pj = proj + (size_t)l * kShadowProj;
    auto mn = [](float a, float b) { return fminf(a, b); };
    auto mx = [](float a, float b) { return fmaxf(a, b); };
    if (kind == kShadowKindPoint) {
        const float3 lp = f3(L[1], L[2], L[3]);
        float far = 0.f;
        for (uint32_t i = threadIdx.x; i < n_inst; i += kSlBlock) {
            const float *b = box + (size_t)i * 6;
            const float3 lo = box_lo(b) - lp, hi = box_hi(b) - lp;
            const float3 c = f3(fmaxf(fabsf(lo.x), fabsf(hi.x)), fmaxf(fabsf(lo.y), fabsf(hi.y)),
                                fmaxf(fabsf(lo.z), fabsf(hi.z)));
            far = fmaxf(far, sqrtf(dot3(c, c)));
        }
        far = block_reduce(far, sh, mx);
        if (threadIdx.x == 0) {
            const float range = L[10];
            const float cut = range > 0.f ? range * (1.f + kRangeSlack) : INFINITY;
            pj[0] = make_float4(kind, kNearScale * far, cut, 0.f);
            pj[1] = make_float4(lp.x, lp.y, lp.z, 0.f);
            pj[2] = pj[3] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        return;
    }
    if (kind == kShadowKindDir) {
        float3 e1, e2;
        dir_basis(L, e1, e2);
        float u0 = INFINITY, u1 = -INFINITY, v0 = INFINITY, v1 = -INFINITY, big = 0.f;
        for (uint32_t i = threadIdx.x; i < n_inst; i += kSlBlock) {
            const float *b = box + (size_t)i * 6;
            for (int c = 0; c < 8; c++) {
                const float3 p = f3(b[(c & 1) ? 3 : 0], b[(c & 2) ? 4 : 1], b[(c & 4) ? 5 : 2]);
                const float u = dot3(p, e1), v = dot3(p, e2);
                u0 = fminf(u0, u);
                u1 = fmaxf(u1, u);
                v0 = fminf(v0, v);
                v1 = fmaxf(v1, v);
                big = fmaxf(big, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
            }
        }
        u0 = block_reduce(u0, sh, mn);
        u1 = block_reduce(u1, sh, mx);
        v0 = block_reduce(v0, sh, mn);
        v1 = block_reduce(v1, sh, mx);
        big = block_reduce(big, sh, mx);
        if (threadIdx.x == 0) {
            const float margin = kDirMargin * big + 1e-6f;
            const float n = (float)kShadowDirGrid;
            const float lu = u0 - margin, lv = v0 - margin;
            // no box: an empty grid (su = sv = 0 sends every ray to cell 0, whose list is empty)
            const bool any = u0 <= u1;
            pj[0] = make_float4(kind, 0.f, 0.f, 0.f);
            pj[1] = make_float4(e1.x, e1.y, e1.z, any ? lu : 0.f);
            pj[2] = make_float4(e2.x, e2.y, e2.z, any ? lv : 0.f);
            pj[3] = make_float4(any ? n / (u1 + margin - lu) : 0.f, any ? n / (v1 + margin - lv) : 0.f, margin, 0.f);
        }
        return;
    }
    if (threadIdx.x == 0) {
        pj[0] = make_float4(kShadowKindAll, 0.f, 0.f, 0.f);
        pj[1] = pj[2] = pj[3] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// cell range [c0, c1] of a face coordinate range [lo, hi] (already inside [-1, 1])
__device__ __forceinline__ void face_cells(float lo, float hi, uint32_t &c0, uint32_t &c1) {
    const float h = 0.5f * (float)kShadowFace, top = (float)(kShadowFace - 1);
    c0 = (uint32_t)fminf(fmaxf(floorf((lo + 1.f) * h), 0.f), top);
    c1 = (uint32_t)fminf(fmaxf(floorf((hi + 1.f) * h), 0.f), top);
}

// range of y / x over the box part with x > 0 (x in [xl, x1], xl >= 0, x1 > 0; y in
// [y0, y1]): the extremes sit at corners; xl = 0 makes a side unbounded
__device__ __forceinline__ void ratio_range(float y0, float y1, float xl, float x1, float &lo, float &hi) {
    lo = y0 < 0.f ? y0 / xl : y0 / x1;
    hi = y1 > 0.f ? y1 / xl : y1 / x1;
}

// rects[(l * n + i) * 6 + face]: cell rectangle of slot i's box on each face of light l
__global__ void __launch_bounds__(kSlBlock) k_sl_rects(const float4 *__restrict__ proj, const float *__restrict__ box,
                                                       uint32_t n_inst, uint32_t n_lights, uint32_t *__restrict__ rects) {
    const uint32_t gid = blockIdx.x * kSlBlock + threadIdx.x;
    if (gid >= n_inst * n_lights) return;
    const uint32_t l = gid / n_inst, i = gid - l * n_inst;
    const float4 *pj = proj + (size_t)l * kShadowProj;
    const float4 h = pj[0];
    const float *b = box + (size_t)i * 6;
    uint32_t *out = rects + (size_t)gid * 6;
    if (h.x == kShadowKindPoint) {
        const float3 lp = xyz(pj[1]);
        const float3 lo = box_lo(b) - lp, hi = box_hi(b) - lp;
        const float3 gap = f3(fmaxf(fmaxf(lo.x, -hi.x), 0.f), fmaxf(fmaxf(lo.y, -hi.y), 0.f),
                              fmaxf(fmaxf(lo.z, -hi.z), 0.f));
        const float dist = sqrtf(dot3(gap, gap));
        if (dist > h.z + h.y) {  // beyond a ranged light's reach
            for (int f = 0; f < 6; f++) out[f] = kEmptyRect;
            return;
        }
        if (dist <= h.y) {  // at or near the light: every cell
            for (int f = 0; f < 6; f++) out[f] = kFullFace;
            return;
        }
        for (int f = 0; f < 6; f++) {
            const int k = f >> 1;
            const bool neg = f & 1;
            // coordinate along the face's axis, positive toward the face
            const float x0 = neg ? -comp3(hi, k) : comp3(lo, k), x1 = neg ? -comp3(lo, k) : comp3(hi, k);
            if (!(x1 > 0.f)) {
                out[f] = kEmptyRect;
                continue;
            }
            const float xl = fmaxf(x0, 0.f);
            const int ka = k == 0 ? 1 : 0, kb = k == 2 ? 1 : 2;
            float ua, ub, va, vb;
            ratio_range(comp3(lo, ka), comp3(hi, ka), xl, x1, ua, ub);
            ratio_range(comp3(lo, kb), comp3(hi, kb), xl, x1, va, vb);
            ua = fmaxf(ua - kFaceMargin, -1.f);
            ub = fminf(ub + kFaceMargin, 1.f);
            va = fmaxf(va - kFaceMargin, -1.f);
            vb = fminf(vb + kFaceMargin, 1.f);
            if (!(ua <= ub && va <= vb)) {
                out[f] = kEmptyRect;
                continue;
            }
            uint32_t u0, u1, v0, v1;
            face_cells(ua, ub, u0, u1);
            face_cells(va, vb, v0, v1);
            out[f] = pack_rect(u0, u1, v0, v1);
        }
        return;
    }
    for (int f = 1; f < 6; f++) out[f] = kEmptyRect;
    if (h.x == kShadowKindDir) {
        const float4 e1 = pj[1], e2 = pj[2], g = pj[3];
        float u0 = INFINITY, u1 = -INFINITY, v0 = INFINITY, v1 = -INFINITY;
        for (int c = 0; c < 8; c++) {
            const float3 p = f3(b[(c & 1) ? 3 : 0], b[(c & 2) ? 4 : 1], b[(c & 4) ? 5 : 2]);
            const float u = dot3(p, xyz(e1)), v = dot3(p, xyz(e2));
            u0 = fminf(u0, u);
            u1 = fmaxf(u1, u);
            v0 = fminf(v0, v);
            v1 = fmaxf(v1, v);
        }
        const float top = (float)(kShadowDirGrid - 1);
        const float cu0 = fminf(fmaxf(floorf((u0 - g.z - e1.w) * g.x), 0.f), top);
        const float cu1 = fminf(fmaxf(floorf((u1 + g.z - e1.w) * g.x), 0.f), top);
        const float cv0 = fminf(fmaxf(floorf((v0 - g.z - e2.w) * g.y), 0.f), top);
        const float cv1 = fminf(fmaxf(floorf((v1 + g.z - e2.w) * g.y), 0.f), top);
        out[0] = pack_rect((uint32_t)cu0, (uint32_t)cu1, (uint32_t)cv0, (uint32_t)cv1);
        return;
    }
    out[0] = kEmptyRect;
}

// face and (u, v) of cell c of a light of the given kind; false when the cell is unused
__device__ __forceinline__ bool cell_coords(float kind, uint32_t c, uint32_t &face, uint32_t &u, uint32_t &v) {
    if (kind == kShadowKindPoint) {
        face = c / (kShadowFace * kShadowFace);
        v = (c / kShadowFace) % kShadowFace;
        u = c % kShadowFace;
        return true;
    }
    if (kind == kShadowKindDir && c < kShadowDirGrid * kShadowDirGrid) {
        face = 0;
        v = c / kShadowDirGrid;
        u = c % kShadowDirGrid;
        return true;
    }
    return false;
}

// exclusive scan of one value per thread across the block (kSlBlock threads); *sum = the
// block's total
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh, uint32_t *sum) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t d = 1; d < kSlBlock; d <<= 1) {  // inclusive Hillis-Steele scan
        const uint32_t t = threadIdx.x >= d ? sh[threadIdx.x - d] : 0u;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    const uint32_t incl = sh[threadIdx.x];
    *sum = sh[kSlBlock - 1];
    __syncthreads();
    return incl - v;
}

// per (light, cell): offsets[cell] = the cell's offset within its block of kSlBlock cells,
// bsum[block] = the block's entries
__global__ void __launch_bounds__(kSlBlock) k_sl_count(const float4 *__restrict__ proj, const uint32_t *__restrict__ rects,
                                                       uint32_t n_inst, uint32_t n_lights, uint32_t *__restrict__ offsets,
                                                       uint32_t *__restrict__ bsum) {
    __shared__ uint32_t sh[kSlBlock];
    const uint32_t gid = blockIdx.x * kSlBlock + threadIdx.x;
    uint32_t n = 0;
    if (gid < n_lights * kShadowCells) {
        const uint32_t l = gid / kShadowCells, c = gid - l * kShadowCells;
        uint32_t face, u, v;
        if (cell_coords(proj[(size_t)l * kShadowProj].x, c, face, u, v)) {
            const uint32_t *r = rects + (size_t)l * n_inst * 6 + face;
            for (uint32_t i = 0; i < n_inst; i++) n += in_rect(r[(size_t)i * 6], u, v) ? 1u : 0u;
        }
    }
    uint32_t total;
    const uint32_t off = block_exclusive_scan(n, sh, &total);
    if (gid < n_lights * kShadowCells) offsets[gid] = off;
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// exclusive offsets of the n block sums in place, starting at `base`; total[0] = base + sum
__global__ void __launch_bounds__(kSlBlock) k_sl_scan(uint32_t *__restrict__ bsum, uint32_t n, uint32_t base,
                                                      uint32_t *__restrict__ total) {
    __shared__ uint32_t sh[kSlBlock];
    uint32_t run = base;
    for (uint32_t c0 = 0; c0 < n; c0 += kSlBlock) {  // chunks of kSlBlock sums
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t v = i < n ? bsum[i] : 0u;
        uint32_t sum;
        const uint32_t off = block_exclusive_scan(v, sh, &sum);
        if (i < n) bsum[i] = run + off;
        run += sum;
    }
    if (threadIdx.x == 0) total[0] = run;
}

// boxes[0, n) = the padded world boxes, boxes[n] = NaN (no slab test passes it)
__global__ void __launch_bounds__(kSlBlock) k_sl_boxes(const float *__restrict__ box, uint32_t n_inst,
                                                       float *__restrict__ boxes) {
    const uint32_t gid = blockIdx.x * kSlBlock + threadIdx.x;
    if (gid < n_inst * 6) boxes[gid] = box[gid];
    else if (gid < (n_inst + 1) * 6) boxes[gid] = __builtin_nanf("");
}

// Sort key of a slot's box for light record pj (lower = walked first): the squared distance
// from a point or spot light to the box, or minus the box centre's height along the
// direction toward a directional light (w = e1 x e2); NaN sorts last.
__device__ __forceinline__ float order_key(const float4 *pj, const float *b) {
    float k = 0.f;
    if (pj[0].x == kShadowKindPoint) {
        const float3 lp = xyz(pj[1]);
        const float dx = fmaxf(fmaxf(b[0] - lp.x, lp.x - b[3]), 0.f);
        const float dy = fmaxf(fmaxf(b[1] - lp.y, lp.y - b[4]), 0.f);
        const float dz = fmaxf(fmaxf(b[2] - lp.z, lp.z - b[5]), 0.f);
        k = dx * dx + dy * dy + dz * dz;
    } else if (pj[0].x == kShadowKindDir) {
        const float3 w = cross3(xyz(pj[1]), xyz(pj[2]));
        k = -(w.x * 0.5f * (b[0] + b[3]) + w.y * 0.5f * (b[1] + b[4]) + w.z * 0.5f * (b[2] + b[5]));
    }
    return k == k ? k : INFINITY;
}

// grid (slot blocks, lights): rank of slot i = how many slots come before it in the total
// order (key, slot); order[l][rank] = i
__global__ void __launch_bounds__(kSlBlock) k_sl_order(const float4 *__restrict__ proj, const float *__restrict__ box,
                                                       uint32_t n_inst, uint16_t *__restrict__ order) {
    __shared__ float keys[kSlBlock];
    const uint32_t l = blockIdx.y, i = blockIdx.x * kSlBlock + threadIdx.x;
    const float4 *pj = proj + (size_t)l * kShadowProj;
    const float ki = i < n_inst ? order_key(pj, box + (size_t)i * 6) : 0.f;
    uint32_t rank = 0;
    for (uint32_t j0 = 0; j0 < n_inst; j0 += kSlBlock) {
        const uint32_t j = j0 + threadIdx.x;
        keys[threadIdx.x] = j < n_inst ? order_key(pj, box + (size_t)j * 6) : INFINITY;
        __syncthreads();
        const uint32_t m = min(kSlBlock, n_inst - j0);
        for (uint32_t t = 0; t < m; t++) {
            const float kj = keys[t];
            rank += (kj < ki || (kj == ki && j0 + t < i)) ? 1u : 0u;
        }
        __syncthreads();
    }
    if (i < n_inst) order[(size_t)l * n_inst + rank] = (uint16_t)i;
}

__global__ void __launch_bounds__(kSlBlock) k_sl_fill(const float4 *__restrict__ proj, const uint32_t *__restrict__ rects,
                                                      const uint32_t *__restrict__ offsets,
                                                      const uint32_t *__restrict__ bsum,
                                                      const uint32_t *__restrict__ total, uint32_t n_inst,
                                                      uint32_t n_lights, uint32_t capacity,
                                                      const uint16_t *__restrict__ order, uint2 *__restrict__ cells,
                                                      uint16_t *__restrict__ list) {
    const uint32_t gid = blockIdx.x * kSlBlock + threadIdx.x;
    if (gid <= n_inst) list[gid] = (uint16_t)gid;  // the identity list [0, n_inst), then the NaN box
    if (gid >= n_lights * kShadowRecs) return;
    const uint32_t l = gid / kShadowRecs, c = gid - l * kShadowRecs;
    const uint2 all = make_uint2(0u, n_inst), none = make_uint2(n_inst, n_inst + 1u);
    if (c == kShadowAll) {
        cells[gid] = all;
        return;
    }
    if (c == kShadowNone) {
        cells[gid] = none;
        return;
    }
    if (total[0] > capacity) {  // lists do not fit: every cell lists every slot
        cells[gid] = all;
        return;
    }
    uint32_t face, u, v, k = 0;
    const uint32_t cell = l * kShadowCells + c;
    const uint32_t first = bsum[cell / kSlBlock] + offsets[cell];
    if (cell_coords(proj[(size_t)l * kShadowProj].x, c, face, u, v)) {
        const uint32_t *r = rects + (size_t)l * n_inst * 6 + face;
        const uint16_t *ord = order + (size_t)l * n_inst;
        for (uint32_t t = 0; t < n_inst; t++) {
            const uint32_t i = ord[t];
            if (in_rect(r[(size_t)i * 6], u, v)) list[first + k++] = (uint16_t)i;
        }
    }
    cells[gid] = k ? make_uint2(first, first + k) : none;
}

}  // namespace

hipError_t launch_shadow_lists(const float *lights, uint32_t n_lights, const float *inst_box, uint32_t n_inst,
                               const ShadowLists &sl, hipStream_t st) {
    if (n_lights == 0 || n_inst == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sl_proj, dim3(n_lights), dim3(kSlBlock), 0, st, lights, inst_box, n_inst, sl.proj);
    const uint32_t pairs = n_lights * n_inst, cells = n_lights * kShadowCells, recs = n_lights * kShadowRecs;
    hipLaunchKernelGGL(k_sl_rects, dim3((pairs + kSlBlock - 1) / kSlBlock), dim3(kSlBlock), 0, st, sl.proj, inst_box,
                       n_inst, n_lights, sl.rects);
    const uint32_t cell_blocks = (cells + kSlBlock - 1) / kSlBlock;
    hipLaunchKernelGGL(k_sl_count, dim3(cell_blocks), dim3(kSlBlock), 0, st, sl.proj, sl.rects, n_inst, n_lights,
                       sl.offsets, sl.bsum);
    const uint32_t nb = (n_inst + 1) * 6;
    hipLaunchKernelGGL(k_sl_boxes, dim3((nb + kSlBlock - 1) / kSlBlock), dim3(kSlBlock), 0, st, inst_box, n_inst,
                       sl.boxes);
    hipLaunchKernelGGL(k_sl_scan, dim3(1), dim3(kSlBlock), 0, st, sl.bsum, cell_blocks, n_inst + 1u, sl.total);
    hipLaunchKernelGGL(k_sl_order, dim3((n_inst + kSlBlock - 1) / kSlBlock, n_lights), dim3(kSlBlock), 0, st, sl.proj,
                       inst_box, n_inst, sl.order);
    const uint32_t fill = std::max(recs, n_inst + 1u);
    hipLaunchKernelGGL(k_sl_fill, dim3((fill + kSlBlock - 1) / kSlBlock), dim3(kSlBlock), 0, st, sl.proj, sl.rects,
                       sl.offsets, sl.bsum, sl.total, n_inst, n_lights, sl.capacity, sl.order, sl.cells, sl.list);
    return hipGetLastError();
}
