// traverse.h — two-level closest-hit and any-hit walks over the device layouts.
//
// Both walks restate traverseClosest/blasClosest and traverseAny/blasAny
// (js/spectral_traverse.js:354-569): threaded pre-order nodes, descend to i+1 on a box
// hit, follow the miss link otherwise, strict '<' on t so the first-visited hit wins
// exact ties. They are written as ONE loop over both levels ("if-if" form): every
// iteration takes exactly one step — a BLAS node of the current instance, the entry
// into the next instance of the current TLAS leaf, or a TLAS node — so the lanes of a
// wave stay in the same loop however deep each one is. The visit order is the
// reference's (a TLAS leaf's instances are walked to completion before the next TLAS
// node). Iterations are bounded, so a corrupt layout cannot hang a wave.
//
// LDS=true reads the TLAS nodes and instance records from the block's LDS copy
// (g_lds: 2 uint4 per TLAS node followed by 4 uint4 per instance).
#pragma once
#include "sbgi_dev.h"

extern __shared__ uint4 g_lds[];

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

template <bool LDS>
__device__ __forceinline__ void tlas_node(const DevScene &s, uint32_t n, uint4 &a, uint4 &b) {
    if (LDS) {
        a = g_lds[(size_t)n * 2];
        b = g_lds[(size_t)n * 2 + 1];
    } else {
        a = s.tlas[(size_t)n * 2];
        b = s.tlas[(size_t)n * 2 + 1];
    }
}

__device__ __forceinline__ float4 as_f4(uint4 u) {
    return make_float4(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z), __uint_as_float(u.w));
}

template <bool LDS>
__device__ __forceinline__ void inst_rows(const DevScene &s, uint32_t inst, float4 &r0, float4 &r1, float4 &r2,
                                          float4 &r3) {
    if (LDS) {
        const uint4 *p = g_lds + (size_t)s.tlas_count * 2 + (size_t)inst * 4;
        r0 = as_f4(p[0]);
        r1 = as_f4(p[1]);
        r2 = as_f4(p[2]);
        r3 = as_f4(p[3]);
    } else {
        const float4 *p = s.inst + (size_t)inst * 4;
        r0 = p[0];
        r1 = p[1];
        r2 = p[2];
        r3 = p[3];
    }
}

// instInvPoint / instInvDir (js/spectral_traverse.js:148-153), evaluated without FMA
// contraction so the local ray is the one the reference's op order produces.
__device__ __forceinline__ float3 inst_point(float4 r0, float4 r1, float4 r2, float3 p) {
#pragma clang fp contract(off)
    return f3(dot3(xyz(r0), p) + r0.w, dot3(xyz(r1), p) + r1.w, dot3(xyz(r2), p) + r2.w);
}
__device__ __forceinline__ float3 inst_dir(float4 r0, float4 r1, float4 r2, float3 d) {
#pragma clang fp contract(off)
    return f3(dot3(xyz(r0), d), dot3(xyz(r1), d), dot3(xyz(r2), d));
}

__device__ __forceinline__ float3 safe_inv(float3 d) { return f3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z); }

// Which node copy a local ray walks: with 8 octant copies, copy o has every interior
// node's children in near-first order for directions whose sign bits are o.
__device__ __forceinline__ uint32_t octant_of(float3 rd, uint32_t octants) {
    if (octants < 8) return 0u;
    return (rd.x < 0.f ? 1u : 0u) | (rd.y < 0.f ? 2u : 0u) | (rd.z < 0.f ? 4u : 0u);
}

__device__ __forceinline__ uint64_t walk_budget(const DevScene &s) {
    return (uint64_t)s.tlas_count * 2 + (uint64_t)s.node_count * 256 + 1024;
}

// Closest hit: traverseClosest + blasClosest (js/spectral_traverse.js:354-474).
template <bool COUNT, bool LDS>
__device__ __forceinline__ void trace_closest(const DevScene &s, float3 wro, float3 wrd, float &best_t, int &best_tri,
                                              int &best_inst, TravCounters<COUNT> &cnt) {
    const float3 winvd = safe_inv(wrd);
    best_t = SBGI_T_MAX;
    best_tri = -1;
    best_inst = -1;
    uint32_t tcur = 0, bcur = 0, bend = 0, inext = 0, iend = 0, cur_inst = 0;
    const uint4 *blas = s.blas;
    float3 ro = wro, rd = wrd, invd = winvd;
    float dsign = 1.f;
    const uint64_t budget = walk_budget(s);
    for (uint64_t it = 0; it < budget; ++it) {
        if (bcur < bend) {
            const uint4 a = blas[(size_t)bcur * 2], b = blas[(size_t)bcur * 2 + 1];
            const float3 bmin = f3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
            const float3 bmax = f3(__uint_as_float(a.w), __uint_as_float(b.x), __uint_as_float(b.y));
            cnt.add_blas();
            float tn, tf;
            slab(bmin, bmax, ro, invd, tn, tf);
            uint32_t next = b.z;
            if (tf >= fmaxf(tn, 0.f) && tn < best_t) {
                if (b.w == 0xFFFFFFFFu) {
                    next = bcur + 1;
                } else {
                    const uint32_t off = b.w & 0x00FFFFFFu, n = b.w >> 24;
                    for (uint32_t i = 0; i < n; i++) {
                        const float4 *tp = s.tri + (size_t)(off + i) * 3;
                        const float4 A = tp[0], B = tp[1], C = tp[2];
                        cnt.add_tri();
                        float th, det;
                        if (mt_test(A, B, C, ro, rd, th, det) && th > SBGI_RAY_EPS && th < best_t) {
                            const uint32_t flags = __float_as_uint(A.w);
                            if (side_ok(flags, det * dsign) && (flags & 4u)) {
                                best_t = th;
                                best_tri = (int)(off + i);
                                best_inst = (int)cur_inst;
                            }
                        }
                    }
                }
            }
            bcur = next;
            continue;
        }
        if (inext < iend) {
            const uint32_t inst = inext++;
            cnt.add_inst();
            float4 r0, r1, r2, r3;
            inst_rows<LDS>(s, inst, r0, r1, r2, r3);
            if (r3.w > 0.5f) {
                ro = inst_point(r0, r1, r2, wro);
                rd = inst_dir(r0, r1, r2, wrd);
                invd = safe_inv(rd);
                dsign = r3.z;
                bcur = (uint32_t)r3.x;
                bend = (uint32_t)r3.y;
                cur_inst = inst;
                blas = s.blas + (size_t)octant_of(rd, s.blas_octants) * s.node_count * 2;
            }
            continue;
        }
        if (tcur >= s.tlas_count) break;
        uint4 a, b;
        tlas_node<LDS>(s, tcur, a, b);
        const float3 bmin = f3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
        const float3 bmax = f3(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z));
        cnt.add_tlas();
        float tn, tf;
        slab(bmin, bmax, wro, winvd, tn, tf);
        if (tf >= fmaxf(tn, 0.f) && tn < best_t) {
            if (b.w == 0xFFFFFFFFu) {
                tcur = tcur + 1;
                continue;
            }
            inext = b.w & 0x00FFFFFFu;
            iend = inext + (b.w >> 24);
        }
        tcur = a.w;
    }
}

// Any hit within (RAY_EPS, maxd): traverseAny + blasAny (js/spectral_traverse.js:477-569),
// ending as soon as `blocked` is set (the reference's blockedVar/Break). The result does
// not depend on visit order.
template <bool COUNT, bool LDS>
__device__ __forceinline__ bool trace_any(const DevScene &s, float3 wro, float3 wrd, float maxd,
                                          TravCounters<COUNT> &cnt) {
    const float3 winvd = safe_inv(wrd);
    uint32_t tcur = 0, bcur = 0, bend = 0, inext = 0, iend = 0;
    const uint4 *blas = s.blas;
    float3 ro = wro, rd = wrd, invd = winvd;
    float dsign = 1.f;
    bool blocked = false;
    const uint64_t budget = walk_budget(s);
    for (uint64_t it = 0; it < budget && !blocked; ++it) {
        if (bcur < bend) {
            const uint4 a = blas[(size_t)bcur * 2], b = blas[(size_t)bcur * 2 + 1];
            const float3 bmin = f3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
            const float3 bmax = f3(__uint_as_float(a.w), __uint_as_float(b.x), __uint_as_float(b.y));
            cnt.add_blas();
            float tn, tf;
            slab(bmin, bmax, ro, invd, tn, tf);
            uint32_t next = b.z;
            if (tf >= fmaxf(tn, 0.f) && tn < maxd) {
                if (b.w == 0xFFFFFFFFu) {
                    next = bcur + 1;
                } else {
                    const uint32_t off = b.w & 0x00FFFFFFu, n = b.w >> 24;
                    for (uint32_t i = 0; i < n && !blocked; i++) {
                        const float4 *tp = s.tri + (size_t)(off + i) * 3;
                        const float4 A = tp[0], B = tp[1], C = tp[2];
                        cnt.add_tri();
                        float th, det;
                        if (mt_test(A, B, C, ro, rd, th, det) && th > SBGI_RAY_EPS && th < maxd) {
                            const uint32_t flags = __float_as_uint(A.w);
                            blocked = side_ok(flags, det * dsign) && (flags & 8u);
                        }
                    }
                }
            }
            bcur = next;
            continue;
        }
        if (inext < iend) {
            const uint32_t inst = inext++;
            cnt.add_inst();
            float4 r0, r1, r2, r3;
            inst_rows<LDS>(s, inst, r0, r1, r2, r3);
            if (r3.w > 0.5f) {
                ro = inst_point(r0, r1, r2, wro);
                rd = inst_dir(r0, r1, r2, wrd);
                invd = safe_inv(rd);
                dsign = r3.z;
                bcur = (uint32_t)r3.x;
                bend = (uint32_t)r3.y;
                blas = s.blas + (size_t)octant_of(rd, s.blas_octants) * s.node_count * 2;
            }
            continue;
        }
        if (tcur >= s.tlas_count) break;
        uint4 a, b;
        tlas_node<LDS>(s, tcur, a, b);
        const float3 bmin = f3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
        const float3 bmax = f3(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z));
        cnt.add_tlas();
        float tn, tf;
        slab(bmin, bmax, wro, winvd, tn, tf);
        if (tf >= fmaxf(tn, 0.f) && tn < maxd) {
            if (b.w == 0xFFFFFFFFu) {
                tcur = tcur + 1;
                continue;
            }
            inext = b.w & 0x00FFFFFFu;
            iend = inext + (b.w >> 24);
        }
        tcur = a.w;
    }
    return blocked;
}

// Copy the TLAS nodes and instance records into the block's LDS (LDS=true walks).
__device__ __forceinline__ void stage_tlas_lds(const DevScene &s) {
    const uint32_t nt = s.tlas_count * 2, ni = s.inst_count * 4;
    const float4 *inst = s.inst;
    for (uint32_t i = threadIdx.x; i < nt + ni; i += blockDim.x) {
        if (i < nt) {
            g_lds[i] = s.tlas[i];
        } else {
            const float4 v = inst[i - nt];
            g_lds[i] = make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w));
        }
    }
    __syncthreads();
}
