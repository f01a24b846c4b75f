// traverse.h — shared pieces of the two-level walk: traversal counters, TLAS node and
// instance-record access (LDS or global), the instance transform of the ray, the node
// copy a ray walks, and the LDS staging of the TLAS.
//
// LDS=true reads the TLAS nodes and instance records from the block's LDS copy
// (g_lds: 2 uint4 per TLAS node followed by 4 uint4 per instance).
#pragma once
#include "sbgi_dev.h"

extern __shared__ uint4 g_lds[];

template <bool COUNT>
struct TravCounters {
    uint32_t blas = 0, tlas = 0, tris = 0, insts = 0;
    // wave steps of the NODE, TRI and INST phases and of refills (counted by lane 0)
    uint32_t node_w = 0, tri_w = 0, inst_w = 0, refill_w = 0;
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
        const uint4 *p = g_lds + (size_t)s.lds_inst + (size_t)inst * 4;
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

// Which node copy a local ray walks: with 8 octant copies, copy o has every interior
// node's children in near-first order for directions whose sign bits are o.
__device__ __forceinline__ uint32_t octant_of(float3 rd, uint32_t octants) {
    if (octants < 8) return 0u;
    return (rd.x < 0.f ? 1u : 0u) | (rd.y < 0.f ? 2u : 0u) | (rd.z < 0.f ? 4u : 0u);
}

// Box of instance slot `slot` as a one-instance TLAS leaf of a candidate list (the
// shadow walk's lists, k_shadow.hip): {min.xyz, -}, {max.xyz, -} from the block's LDS copy
// (after the instance records) or from the list boxes (the padded world box of every slot,
// then at slot = instances a NaN box, which no ray hits).
template <bool LDS>
__device__ __forceinline__ void list_box(const DevScene &s, uint32_t slot, uint4 &a, uint4 &b) {
    if (LDS) {
        a = g_lds[(size_t)s.lds_box + (size_t)slot * 2];
        b = g_lds[(size_t)s.lds_box + (size_t)slot * 2 + 1];
    } else {
        const float *p = s.inst_box + (size_t)slot * 6;
        a = make_uint4(__float_as_uint(p[0]), __float_as_uint(p[1]), __float_as_uint(p[2]), 0u);
        b = make_uint4(__float_as_uint(p[3]), __float_as_uint(p[4]), __float_as_uint(p[5]), 0u);
    }
}

// Copy the instance records and the list boxes (one per slot, then the never-hit box)
// into the block's LDS (walks of candidate lists): records at lds_inst, boxes at lds_box.
__device__ __forceinline__ void stage_list_lds(const DevScene &s) {
    const uint32_t ni = s.inst_count * 4, nb = (s.inst_count + 1) * 2;
    for (uint32_t i = threadIdx.x; i < ni + nb; i += blockDim.x) {
        if (i < ni) {
            const float4 v = s.inst[i];
            g_lds[s.lds_inst + i] = make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z),
                                               __float_as_uint(v.w));
        } else {
            const uint32_t j = i - ni, slot = j >> 1, hi = j & 1u;
            const float *p = s.inst_box + (size_t)slot * 6 + hi * 3;
            g_lds[s.lds_box + j] = make_uint4(__float_as_uint(p[0]), __float_as_uint(p[1]), __float_as_uint(p[2]), 0u);
        }
    }
    __syncthreads();
}

// Copy the TLAS nodes (unless lds_inst is 0) and the instance records into the block's
// LDS (LDS=true walks).
__device__ __forceinline__ void stage_tlas_lds(const DevScene &s) {
    const uint32_t nt = s.lds_inst, ni = s.inst_count * 4;
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
