// sbgi_dev.h — device layouts and traversal routines for gfx950.
//
// Device-side scene layouts (derived on the device from the reference buffers by the
// pack kernels in k_scene.hip):
//   blas  uint4[2*N]   the reference 32-byte node unchanged: {min.xyz, max.x}, {max.yz, miss,
//                      payload} — two dwordx4 loads per node.
//   tri   float4[3*T]  {p0.xyz, flags}, {e1.xyz, material}, {e2.xyz, 0}: one 48-byte record
//                      replaces triIndex + 3 scattered vertices + triMaterial + material flags.
//                      flags: bits 0-1 side (0 front, 1 back, 2 double), bit 2 closest-hit
//                      accept, bit 3 shadow-ray accept (materialTraversalFlags,
//                      js/spectral_traverse.js:193-220, 410-423, 512-516).
//   tlas  uint4[2*M]   {min.xyz, miss}, {max.xyz, leaf} with leaf = 0xFFFFFFFF for interior,
//                      else (count << 24) | instance slot offset.
//   inst  float4[4*I]  inverse world rows 0..2 (w = translation), {blasRoot, blasEnd,
//                      detSign, active} (js/spectral_traverse.js:140-158).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sbgi.h"

#define SBGI_RAY_EPS 1e-3f
#define SBGI_T_MAX 1e30f
#define SBGI_PI 3.141592653589793f

// Offset of a hit along its normal, and so of its shadow rays' origin and extent: 0.5 % of
// the cascade's smallest cell, at least the ray epsilon (js/gi_probes.js:1993-2054). Host
// and device evaluate it alike (two products and a max).
__host__ __device__ __forceinline__ float hit_bias(float cell_min, float trace_bias, float ray_eps) {
    return fmaxf(cell_min * 0.005f * trace_bias, SBGI_RAY_EPS * ray_eps);
}
#define SBGI_GOLDEN_ANGLE 2.39996322972865332f
#define SBGI_TWO_PI 6.283185307179586f

struct DevScene {
    const uint4 *__restrict__ blas;
    const uint4 *__restrict__ nodes;  // BLAS in the TLAS node format {min.xyz, miss}, {max.xyz, payload}
    const float4 *__restrict__ tri;
    const uint4 *__restrict__ tlas;
    // near-first copies of the TLAS, one per ray direction octant (tlas_count nodes each;
    // null: the closest-hit walk reads the single TLAS)
    const uint4 *__restrict__ tlas_oct;
    const float4 *__restrict__ inst;
    const float *__restrict__ mats;
    const uint32_t *__restrict__ tri_mat;
    const float *__restrict__ lights;
    uint32_t node_count, tlas_count, light_count, tri_count, inst_count;
    uint32_t blas_octants;  // 1: reference node order only; 8: + near-first copy per ray octant
    uint32_t shadow_octants;  // copies the shadow walks take: blas_octants, or 1 (one copy for all)
    // LDS walks: uint4 offset of the instance records in the block's LDS (after the TLAS
    // nodes, or 0 when the walk reads its TLAS from global memory and stages only them)
    uint32_t lds_inst;
    // differentiable mode (null otherwise): per-instance-triangle albedo overriding the
    // material colour, albedo[(map[inst].x + tri - map[inst].y) * 3 + c], and a scale per
    // light on its colour x intensity
    const float *__restrict__ albedo;
    const int2 *__restrict__ albedo_map;
    const float *__restrict__ light_scale;
    // wave schedule of the walk (traverse2.h): a TRI / INST phase runs when the lanes
    // waiting in it are >= 1/tri_ratio (1/inst_ratio) of those still in NODE; retired
    // lanes are refilled in batches of >= refill_min; fetch_chunk rays per fetch atomic
    uint32_t tri_ratio, inst_ratio, refill_min, fetch_chunk;
    // clustered NEE (null otherwise): per-cell light lists {count, ids[16]} over the cells
    // of cascade 0's probe grid, lg_cells = res - 1 per axis, spanning lg_min + lg_size
    const uint32_t *__restrict__ light_grid;
    float lg_min[3], lg_size[3];
    uint32_t lg_cells[3];
    // emitter records (type 3) injected as virtual rays: lights[emitter_base + e]
    uint32_t emitter_base, emitter_count;
    // SBGI_TRAV_WATERTIGHT: the walks test triangles with wt_test on the stored vertices
    // (tri_index into vdata, 8 floats per vertex) instead of Moller-Trumbore on tri
    uint32_t watertight;
    const uint32_t *__restrict__ tri_index;
    const float *__restrict__ vdata;
    // light-space candidate lists of the shadow-ray walk (k_shadow.hip; null: the shadow
    // walk walks the TLAS): per light kShadowProj float4 of projection, kShadowRecs cell
    // records {first, end} into sl_list (instance slots), and the boxes the list steps test
    // in place of TLAS nodes (6 floats: the padded world box of every slot, then at slot =
    // inst_count a NaN box that no ray hits, the one entry of a cell no box reaches)
    const float4 *__restrict__ sl_proj;
    const uint2 *__restrict__ sl_cells;
    const uint16_t *__restrict__ sl_list;
    const float *__restrict__ inst_box;
    // LDS walks of the lists: uint4 offset of the boxes (2 per slot, after the records)
    uint32_t lds_box;
};

// Light-space candidate lists. A shadow ray runs from a hit toward its light, so every
// point of it lies in the cone of directions from the light that contains the hit (point
// and spot lights), or on the line through the hit along the light direction (directional
// lights). Cells of those directions (a cube map of kShadowFace^2 cells per face around the
// light) or of that plane (a kShadowDirGrid^2 grid across the light direction) list,
// nearest the light first, the instances whose padded world box reaches the cell; the walk
// tests that list instead of the TLAS.
constexpr uint32_t kShadowFace = 16;
constexpr uint32_t kShadowDirGrid = 32;
constexpr uint32_t kShadowCells = 6 * kShadowFace * kShadowFace;
static_assert(kShadowDirGrid * kShadowDirGrid <= kShadowCells, "directional grid exceeds the cell records");
constexpr uint32_t kShadowAll = kShadowCells;       // record {0, instances}: every slot
constexpr uint32_t kShadowNone = kShadowCells + 1;  // record {instances, instances + 1}: the NaN box
constexpr uint32_t kShadowRecs = kShadowCells + 2;  // cell records per light
constexpr uint32_t kShadowProj = 4;                 // float4 of projection per light
// projection kinds (x of a light's first float4)
constexpr float kShadowKindPoint = 0.f, kShadowKindDir = 1.f, kShadowKindAll = 2.f;

// Cell record index of a shadow ray from o for the light whose projection is pj:
//   point / spot {kind, ...}, {position, 0}: the cube-map cell of o - position (face by the
//     largest |component|, +x -x +y -y +z -z; u, v = the two others over it, in [-1, 1]);
//   directional {kind, ...}, {e1, u0}, {e2, v0}, {su, sv, ...}: the grid cell of
//     ((o.e1 - u0) su, (o.e2 - v0) sv), or kShadowNone off the grid (no box reaches it);
//   anything else: kShadowAll.
// The lists are built with margins that cover this function's rounding.
__device__ __forceinline__ uint32_t shadow_cell(const float4 *__restrict__ pj, float3 o) {
    const float kind = pj[0].x;
    if (kind == kShadowKindPoint) {
        const float4 lp = pj[1];
        const float vx = o.x - lp.x, vy = o.y - lp.y, vz = o.z - lp.z;
        const float ax = fabsf(vx), ay = fabsf(vy), az = fabsf(vz);
        const bool fx = ax >= ay && ax >= az, fy = !fx && ay >= az;
        const float m = fx ? ax : (fy ? ay : az), major = fx ? vx : (fy ? vy : vz);
        const float a = fx ? vy : vx, b = fx || fy ? vz : vy;
        const uint32_t face = (fx ? 0u : (fy ? 2u : 4u)) + (major < 0.f ? 1u : 0u);
        // (a / m + 1) N / 2 with the hardware reciprocal; m = 0 gives NaN, which the clamps
        // send to cell 0
        const float h = 0.5f * (float)kShadowFace, sc = h * __builtin_amdgcn_rcpf(m);
        const float top = (float)(kShadowFace - 1);
        const uint32_t iu = (uint32_t)fminf(fmaxf(a * sc + h, 0.f), top);
        const uint32_t iv = (uint32_t)fminf(fmaxf(b * sc + h, 0.f), top);
        return (face * kShadowFace + iv) * kShadowFace + iu;
    }
    if (kind == kShadowKindDir) {
        const float4 e1 = pj[1], e2 = pj[2], g = pj[3];
        const float fu = (o.x * e1.x + o.y * e1.y + o.z * e1.z - e1.w) * g.x;
        const float fv = (o.x * e2.x + o.y * e2.y + o.z * e2.z - e2.w) * g.y;
        const float n = (float)kShadowDirGrid;
        if (!(fu >= 0.f && fu < n && fv >= 0.f && fv < n)) return kShadowNone;
        return (uint32_t)fv * kShadowDirGrid + (uint32_t)fu;
    }
    return kShadowAll;
}

// albedo entry of (instance slot, pool triangle) in the differentiable mode
__device__ __forceinline__ uint32_t albedo_index(const DevScene &s, uint32_t inst, uint32_t tri) {
    const int2 m = s.albedo_map[inst];
    return (uint32_t)(m.x + ((int)tri - m.y));
}

// Per-solve cascade geometry (per-cascade uniforms of js/gi_probes.js:3200-3211).
struct DevCascade {
    float gmin[3], gsize[3];
    uint32_t res[3];
    uint32_t total;       // probeTotal
    uint32_t offset;      // probeOffset
    uint32_t updated;     // updatedCount
    uint32_t rays;        // R
    uint32_t depth_oct;   // 6 or 14
    float max_dist, cell_min;
    uint32_t aw, ah;      // irradiance atlas size in texels
    uint32_t daw, dah;    // depth atlas size
    uint32_t rough;       // reflection lobes on: unshaded hits record -(t + 2) as their distance
};

// Fetch counters of the persistent walks: kFetchHeads per walk, kHeadStride uint32 words
// (one 128-byte line) apart.
constexpr uint32_t kFetchHeads = 64;
constexpr uint32_t kHeadStride = 32;
// default wave schedule of the walks (DevScene::tri_ratio ... fetch_chunk)
constexpr uint32_t kDefaultTriRatio = 2, kDefaultInstRatio = 4, kDefaultRefillMin = 16, kDefaultFetchChunk = 64;

// Traversal counters of counting ticks: [0..3] closest-hit walk {BLAS nodes, TLAS nodes,
// triangle tests, instance entries}, [4] shadow rays emitted, [5..8] the shadow walk's
// {BLAS nodes, TLAS nodes, triangle tests, instance entries}.
// walk[0..3]: BLAS nodes, TLAS nodes, triangle tests, instance entries (lane visits);
// walk[4..7]: wave steps of the NODE, TRI, INST phases and refills; same for shadow_walk
struct Counters {
    unsigned long long walk[8], shadow_rays, shadow_walk[8];
};

__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 operator*(float3 a, float3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 normalize3(float3 v) {
    float len = sqrtf(dot3(v, v));
    return f3(v.x / len, v.y / len, v.z / len);
}
__device__ __forceinline__ float3 xyz(float4 v) { return f3(v.x, v.y, v.z); }

// slab test (js/spectral_traverse.js:372-377)
__device__ __forceinline__ void slab(float3 bmin, float3 bmax, float3 ro, float3 invd, float &tn, float &tf) {
    float t0x = (bmin.x - ro.x) * invd.x, t1x = (bmax.x - ro.x) * invd.x;
    float t0y = (bmin.y - ro.y) * invd.y, t1y = (bmax.y - ro.y) * invd.y;
    float t0z = (bmin.z - ro.z) * invd.z, t1z = (bmax.z - ro.z) * invd.z;
    tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
}

// The same slab test with each plane distance fused, t = fma(b, 1/d, -o/d): one rounding
// instead of two, so a box decision can differ from slab()'s only for a ray that grazes the
// box face. Where 1/d is infinite and b/d and o/d are infinities of one sign the fused form
// gives NaN, which fminf/fmaxf pass over, so the box test only grows more permissive there.
// For any-hit walks, whose answer no extra box visit can change.
__device__ __forceinline__ void slab_fma(float3 bmin, float3 bmax, float3 noid, float3 invd, float &tn, float &tf) {
    const float t0x = fmaf(bmin.x, invd.x, noid.x), t1x = fmaf(bmax.x, invd.x, noid.x);
    const float t0y = fmaf(bmin.y, invd.y, noid.y), t1y = fmaf(bmax.y, invd.y, noid.y);
    const float t0z = fmaf(bmin.z, invd.z, noid.z), t1z = fmaf(bmax.z, invd.z, noid.z);
    tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
}

__device__ __forceinline__ bool side_ok(uint32_t flags, float det) {
    uint32_t side = flags & 3u;
    return side == 2u || (side == 0u && det > 0.f) || (side == 1u && det < 0.f);
}

// Moller-Trumbore with the reference's tests and epsilon (js/spectral_traverse.js:388-409).
// Evaluated without FMA contraction and with a correctly rounded 1/det, the operation
// order of the reference's graph, so edge decisions match the FP32 restatement.
// RCP = true takes 1/det from the hardware reciprocal (1 ulp) instead: the any-hit walk,
// whose answer can then differ only for a ray grazing a triangle edge.
template <bool RCP = false>
__device__ __forceinline__ bool mt_test(float4 A, float4 B, float4 C, float3 ro, float3 rd, float &t, float &det) {
#pragma clang fp contract(off)
    float3 p0 = xyz(A), e1 = xyz(B), e2 = xyz(C);
    float3 pv = cross3(rd, e2);
    det = dot3(e1, pv);
    if (!(fabsf(det) > 1e-12f)) return false;
    float inv_det = RCP ? __builtin_amdgcn_rcpf(det) : 1.0f / det;
    float3 tv = ro - p0;
    float u = dot3(tv, pv) * inv_det;
    if (!(u >= 0.f && u <= 1.f)) return false;
    float3 qv = cross3(tv, e1);
    float v = dot3(rd, qv) * inv_det;
    if (!(v >= 0.f && u + v <= 1.f)) return false;
    t = dot3(e2, qv) * inv_det;
    return true;
}

// Watertight ray/triangle test (Woop, Benthin and Wald 2013, JCGT 2(1)): the
// SBGI_TRAV_WATERTIGHT mode's triangle test, in the same operation order as the oracle's
// restatement so the two agree bit for bit. The triangle is translated to the ray origin and
// sheared so the ray runs along +z (kz = the axis of the largest |d|, kx and ky the next
// two, swapped when d[kz] < 0 to keep the winding); the 2-D edge functions U, V, W of the
// sheared vertices decide the hit, recomputed in double precision when one is exactly zero,
// so two triangles that share an edge cannot both miss a ray through it. The vertices are
// the stored ones (not p0 + edge), which the shared-edge argument needs.
// det (= U + V + W) has the sign of the Moller-Trumbore determinant, so the side rules read
// it unchanged; t = (U z0 + V z1 + W z2) / det.
struct WtRay {
    int kx, ky, kz;
    float sx, sy, sz;
};
__device__ __forceinline__ float comp(float3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ WtRay wt_setup(float3 rd) {
#pragma clang fp contract(off)
    const float ax = fabsf(rd.x), ay = fabsf(rd.y), az = fabsf(rd.z);
    const int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    if (comp(rd, kz) < 0.f) {
        const int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    WtRay w;
    w.kx = kx;
    w.ky = ky;
    w.kz = kz;
    const float dz = comp(rd, kz);
    w.sx = comp(rd, kx) / dz;
    w.sy = comp(rd, ky) / dz;
    w.sz = 1.0f / dz;
    return w;
}
__device__ __forceinline__ bool wt_test(float3 p0, float3 p1, float3 p2, float3 ro, const WtRay &w, float &t,
                                        float &det) {
#pragma clang fp contract(off)
    const float3 a = p0 - ro, b = p1 - ro, c = p2 - ro;
    const float x0 = comp(a, w.kx) - w.sx * comp(a, w.kz), y0 = comp(a, w.ky) - w.sy * comp(a, w.kz);
    const float x1 = comp(b, w.kx) - w.sx * comp(b, w.kz), y1 = comp(b, w.ky) - w.sy * comp(b, w.kz);
    const float x2 = comp(c, w.kx) - w.sx * comp(c, w.kz), y2 = comp(c, w.ky) - w.sy * comp(c, w.kz);
    float U = x2 * y1 - y2 * x1, V = x0 * y2 - y0 * x2, W = x1 * y0 - y1 * x0;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = (float)((double)x2 * (double)y1 - (double)y2 * (double)x1);
        V = (float)((double)x0 * (double)y2 - (double)y0 * (double)x2);
        W = (float)((double)x1 * (double)y0 - (double)y1 * (double)x0);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    det = U + V + W;
    if (det == 0.f) return false;
    const float z0 = w.sz * comp(a, w.kz), z1 = w.sz * comp(b, w.kz), z2 = w.sz * comp(c, w.kz);
    const float T = U * z0 + V * z1 + W * z2;
    t = T / det;
    return true;
}
