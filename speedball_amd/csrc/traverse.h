// traverse.h — two-level closest-hit and any-hit walks over the device layouts.
//
// The stackless walk restates traverseClosest/blasClosest and traverseAny/blasAny
// (js/spectral_traverse.js:354-569): threaded pre-order nodes, descend to i+1 on a
// box hit, follow the miss link otherwise, strict '<' on t so the first-visited hit
// wins exact ties. Iterations are bounded by the node counts exactly as the reference
// bounds its Loop(), so a corrupt layout cannot hang a wave.
#pragma once
#include "sbgi_dev.h"

template <bool COUNT>
struct TravCounters {
    uint32_t blas = 0, tlas = 0, tris = 0, insts = 0;
    __device__ __forceinline__ void add_blas() {
        if (COUNT) blas++;
    }
    __device__ __forceinline__ void add_tlas() {
        if (COUNT) tlas++;
    }
    __device__ __forceinline__ void add_tri() {
        if (COUNT) tris++;
    }
    __device__ __forceinline__ void add_inst() {
        if (COUNT) insts++;
    }
};

__device__ __forceinline__ void inst_rows(const DevScene &s, uint32_t inst, float4 &r0, float4 &r1, float4 &r2,
                                          float4 &r3) {
    const float4 *p = s.inst + (size_t)inst * 4;
    r0 = p[0];
    r1 = p[1];
    r2 = p[2];
    r3 = p[3];
}

__device__ __forceinline__ float3 inst_point(float4 r0, float4 r1, float4 r2, float3 p) {
    return f3(dot3(xyz(r0), p) + r0.w, dot3(xyz(r1), p) + r1.w, dot3(xyz(r2), p) + r2.w);
}
__device__ __forceinline__ float3 inst_dir(float4 r0, float4 r1, float4 r2, float3 d) {
    return f3(dot3(xyz(r0), d), dot3(xyz(r1), d), dot3(xyz(r2), d));
}

__device__ __forceinline__ float3 safe_inv(float3 d) { return f3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z); }

// blasClosest (js/spectral_traverse.js:354-435)
template <bool COUNT>
__device__ __forceinline__ void blas_closest(const DevScene &s, uint32_t inst, float3 wro, float3 wrd, float &best_t,
                                             int &best_tri, int &best_inst, TravCounters<COUNT> &cnt) {
    float4 r0, r1, r2, r3;
    inst_rows(s, inst, r0, r1, r2, r3);
    float3 ro = inst_point(r0, r1, r2, wro);
    float3 rd = inst_dir(r0, r1, r2, wrd);
    float dsign = r3.z;
    float3 invd = safe_inv(rd);
    uint32_t end = (uint32_t)r3.y;
    uint32_t cur = (uint32_t)r3.x;
    for (uint32_t it = 0; it < s.node_count && cur < end; ++it) {
        uint4 a = s.blas[(size_t)cur * 2], b = s.blas[(size_t)cur * 2 + 1];
        float3 bmin = f3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
        float3 bmax = f3(__uint_as_float(a.w), __uint_as_float(b.x), __uint_as_float(b.y));
        cnt.add_blas();
        float tn, tf;
        slab(bmin, bmax, ro, invd, tn, tf);
        if (tf >= fmaxf(tn, 0.f) && tn < best_t) {
            if (b.w == 0xFFFFFFFFu) {
                cur = cur + 1;
                continue;
            }
            uint32_t off = b.w & 0x00FFFFFFu, n = b.w >> 24;
            for (uint32_t i = 0; i < n; i++) {
                uint32_t tri = off + i;
                const float4 *tp = s.tri + (size_t)tri * 3;
                float4 A = tp[0], B = tp[1], C = tp[2];
                cnt.add_tri();
                float th, det;
                if (!mt_test(A, B, C, ro, rd, th, det)) continue;
                if (!(th > SBGI_RAY_EPS && th < best_t)) continue;
                uint32_t flags = __float_as_uint(A.w);
                if (side_ok(flags, det * dsign) && (flags & 4u)) {
                    best_t = th;
                    best_tri = (int)tri;
                    best_inst = (int)inst;
                }
            }
        }
        cur = b.z;
    }
}

// traverseClosest (js/spectral_traverse.js:441-474)
template <bool COUNT>
__device__ __forceinline__ void trace_closest(const DevScene &s, float3 ro, float3 rd, float &best_t, int &best_tri,
                                              int &best_inst, TravCounters<COUNT> &cnt) {
    float3 invd = safe_inv(rd);
    best_t = SBGI_T_MAX;
    best_tri = -1;
    best_inst = -1;
    uint32_t cur = 0;
    for (uint32_t it = 0; it < s.tlas_count && cur < s.tlas_count; ++it) {
        uint4 a = s.tlas[(size_t)cur * 2], b = s.tlas[(size_t)cur * 2 + 1];
        float3 bmin = f3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
        float3 bmax = f3(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z));
        cnt.add_tlas();
        float tn, tf;
        slab(bmin, bmax, ro, invd, tn, tf);
        if (tf >= fmaxf(tn, 0.f) && tn < best_t) {
            if (b.w == 0xFFFFFFFFu) {
                cur = cur + 1;
                continue;
            }
            uint32_t off = b.w & 0x00FFFFFFu, n = b.w >> 24;
            for (uint32_t i = 0; i < n; i++) {
                uint32_t inst = off + i;
                cnt.add_inst();
                if (s.inst[(size_t)inst * 4 + 3].w > 0.5f)
                    blas_closest<COUNT>(s, inst, ro, rd, best_t, best_tri, best_inst, cnt);
            }
        }
        cur = a.w;
    }
}

// blasAny (js/spectral_traverse.js:477-536)
template <bool COUNT>
__device__ __forceinline__ bool blas_any(const DevScene &s, uint32_t inst, float3 wro, float3 wrd, float maxd,
                                         TravCounters<COUNT> &cnt) {
    float4 r0, r1, r2, r3;
    inst_rows(s, inst, r0, r1, r2, r3);
    float3 ro = inst_point(r0, r1, r2, wro);
    float3 rd = inst_dir(r0, r1, r2, wrd);
    float dsign = r3.z;
    float3 invd = safe_inv(rd);
    uint32_t end = (uint32_t)r3.y;
    uint32_t cur = (uint32_t)r3.x;
    for (uint32_t it = 0; it < s.node_count && cur < end; ++it) {
        uint4 a = s.blas[(size_t)cur * 2], b = s.blas[(size_t)cur * 2 + 1];
        float3 bmin = f3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
        float3 bmax = f3(__uint_as_float(a.w), __uint_as_float(b.x), __uint_as_float(b.y));
        cnt.add_blas();
        float tn, tf;
        slab(bmin, bmax, ro, invd, tn, tf);
        if (tf >= fmaxf(tn, 0.f) && tn < maxd) {
            if (b.w == 0xFFFFFFFFu) {
                cur = cur + 1;
                continue;
            }
            uint32_t off = b.w & 0x00FFFFFFu, n = b.w >> 24;
            for (uint32_t i = 0; i < n; i++) {
                uint32_t tri = off + i;
                const float4 *tp = s.tri + (size_t)tri * 3;
                float4 A = tp[0], B = tp[1], C = tp[2];
                cnt.add_tri();
                float th, det;
                if (!mt_test(A, B, C, ro, rd, th, det)) continue;
                if (!(th > SBGI_RAY_EPS && th < maxd)) continue;
                uint32_t flags = __float_as_uint(A.w);
                if (side_ok(flags, det * dsign) && (flags & 8u)) return true;
            }
        }
        cur = b.z;
    }
    return false;
}

// traverseAny (js/spectral_traverse.js:539-569). The result does not depend on visit
// order, so any walk that visits every candidate gives the reference's answer.
template <bool COUNT>
__device__ __forceinline__ bool trace_any(const DevScene &s, float3 ro, float3 rd, float maxd,
                                          TravCounters<COUNT> &cnt) {
    float3 invd = safe_inv(rd);
    uint32_t cur = 0;
    for (uint32_t it = 0; it < s.tlas_count && cur < s.tlas_count; ++it) {
        uint4 a = s.tlas[(size_t)cur * 2], b = s.tlas[(size_t)cur * 2 + 1];
        float3 bmin = f3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
        float3 bmax = f3(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z));
        cnt.add_tlas();
        float tn, tf;
        slab(bmin, bmax, ro, invd, tn, tf);
        if (tf >= fmaxf(tn, 0.f) && tn < maxd) {
            if (b.w == 0xFFFFFFFFu) {
                cur = cur + 1;
                continue;
            }
            uint32_t off = b.w & 0x00FFFFFFu, n = b.w >> 24;
            for (uint32_t i = 0; i < n; i++) {
                uint32_t inst = off + i;
                cnt.add_inst();
                if (s.inst[(size_t)inst * 4 + 3].w > 0.5f && blas_any<COUNT>(s, inst, ro, rd, maxd, cnt)) return true;
            }
        }
        cur = a.w;
    }
    return false;
}
