// k_scene.hip — device-side derivation of the packed traversal layouts and the refits.
//
// pack_*: turn the reference buffers (bvhNodes/triIndex/vertexData/triMaterial/materials,
// js/spectral_scene.js:1188-1366) into the device layouts described in sbgi_dev.h.
// refit: refitFlatBlasRange (js/gi_refit.js:34-165) as a level-synchronous sweep, one
// launch per node height, and refitFlatTlasRange (js/gi_refit.js:245-292) over the small
// reference TLAS as a climb in which the second child to arrive at a parent (per-node
// arrival counter) computes the union. min/max are exact, so both equal the reference's
// sequential reverse pre-order sweep bit for bit.
#include "kernels.h"

namespace {

// materialTraversalFlags + materialSideAccepts inputs folded into one word per triangle
// (js/spectral_scene.js:285-290, js/spectral_traverse.js:193-220)
__device__ __forceinline__ uint32_t tri_flags(const float *__restrict__ mats, uint32_t m) {
    const float *r = mats + (size_t)m * SBGI_MAT_STRIDE;
    const float side = r[22];
    uint32_t sc = 3u;  // NaN side accepts nothing
    if (side < 0.5f)
        sc = 0u;
    else if (side >= 0.5f && side < 1.5f)
        sc = 1u;
    else if (side >= 1.5f)
        sc = 2u;
    const uint32_t fl = (uint32_t)r[26];
    const float alpha = r[10], at = r[23];
    const bool alpha_ok = alpha > 1.0e-4f && (at <= 0.f || alpha >= at);
    const bool acc_closest = (fl & 1u) ? alpha_ok : ((fl & 2u) != 0u);
    const bool acc_any = (fl & 4u) && acc_closest;
    return sc | (acc_closest ? 4u : 0u) | (acc_any ? 8u : 0u);
}

// All triangles, or (ranges != nullptr) the triangles [ranges[y].x, ranges[y].y) of range
// y = blockIdx.y.
__global__ void k_pack_tris(const uint32_t *__restrict__ tri_index, const float *__restrict__ vdata,
                            const uint32_t *__restrict__ tri_mat, const float *__restrict__ mats, uint32_t tri_count,
                            const uint2 *__restrict__ ranges, float4 *__restrict__ tri) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (ranges) {
        const uint2 r = ranges[blockIdx.y];
        t += r.x;
        if (t >= r.y) return;
    }
    if (t >= tri_count) return;
    const uint32_t i0 = tri_index[(size_t)t * 3], i1 = tri_index[(size_t)t * 3 + 1], i2 = tri_index[(size_t)t * 3 + 2];
    const float *v0 = vdata + (size_t)i0 * 8, *v1 = vdata + (size_t)i1 * 8, *v2 = vdata + (size_t)i2 * 8;
    const uint32_t m = tri_mat[t];
    tri[(size_t)t * 3 + 0] = make_float4(v0[0], v0[1], v0[2], __uint_as_float(tri_flags(mats, m)));
    tri[(size_t)t * 3 + 1] = make_float4(v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2], __uint_as_float(m));
    tri[(size_t)t * 3 + 2] = make_float4(v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2], 0.f);
}

__global__ void k_pack_tlas(const float *__restrict__ mats, uint32_t tlas_base, uint32_t tlas_count,
                            uint4 *__restrict__ tlas) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= tlas_count) return;
    const float *f = mats + tlas_base + (size_t)n * SBGI_TLAS_STRIDE;
    const uint32_t leaf = f[8] < 0.5f ? 0xFFFFFFFFu : (((uint32_t)f[8] << 24) | ((uint32_t)f[7] & 0x00FFFFFFu));
    tlas[(size_t)n * 2] = make_uint4(__float_as_uint(f[0]), __float_as_uint(f[1]), __float_as_uint(f[2]), (uint32_t)f[6]);
    tlas[(size_t)n * 2 + 1] = make_uint4(__float_as_uint(f[3]), __float_as_uint(f[4]), __float_as_uint(f[5]), leaf);
}

__global__ void k_pack_inst(const float *__restrict__ mats, uint32_t inst_base, uint32_t inst_count,
                            float4 *__restrict__ inst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= inst_count) return;
    const float *f = mats + inst_base + (size_t)i * SBGI_MAT_STRIDE;
#pragma unroll
    for (int q = 0; q < 4; q++) inst[(size_t)i * 4 + q] = make_float4(f[q * 4], f[q * 4 + 1], f[q * 4 + 2], f[q * 4 + 3]);
}

__device__ __forceinline__ float load_coherent(const float *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void union_children(float *__restrict__ nf, uint32_t i, const uint32_t *__restrict__ nu,
                                               uint32_t stride, uint32_t miss_word, bool miss_is_float) {
    const size_t l = (size_t)i + 1;
    const uint32_t r = miss_is_float ? (uint32_t)load_coherent(nf + l * stride + miss_word)
                                     : __hip_atomic_load(nu + l * stride + miss_word, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
    float mn[3], mx[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        mn[a] = fminf(load_coherent(nf + l * stride + a), load_coherent(nf + (size_t)r * stride + a));
        mx[a] = fmaxf(load_coherent(nf + l * stride + 3 + a), load_coherent(nf + (size_t)r * stride + 3 + a));
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        nf[(size_t)i * stride + a] = mn[a];
        nf[(size_t)i * stride + 3 + a] = mx[a];
    }
}

// Bottom-up TLAS refit from per-slot instance world bounds (6 floats each).
__global__ void k_refit_tlas(float *__restrict__ mats, uint32_t tlas_base, uint32_t tlas_count,
                             const float *__restrict__ inst_bounds, const int32_t *__restrict__ parent,
                             uint32_t *__restrict__ arrive) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tlas_count) return;
    float *tl = mats + tlas_base;
    float *nd = tl + (size_t)i * SBGI_TLAS_STRIDE;
    const float cnt = nd[8];
    if (cnt < 0.5f) return;
    const uint32_t off = (uint32_t)nd[7], n = (uint32_t)cnt;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t s = off; s < off + n; s++) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = fminf(mn[a], inst_bounds[(size_t)s * 6 + a]);
            mx[a] = fmaxf(mx[a], inst_bounds[(size_t)s * 6 + 3 + a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        nd[a] = mn[a];
        nd[3 + a] = mx[a];
    }
    int32_t p = parent[i];
    while (p >= 0) {
        __threadfence();
        if (atomicAdd(&arrive[p], 1u) == 0u) return;
        __threadfence();
        union_children(tl, (uint32_t)p, nullptr, SBGI_TLAS_STRIDE, 6, true);
        p = parent[p];
    }
}

// Walk-format node copies {min.xyz, miss}, {max.xyz, payload}: re-gather the bounds of
// copy node i from pool node perm[i] after a refit (structure words are unchanged). All
// n_total copy nodes, or (ranges != nullptr) the nodes [r.x, r.y) of range r =
// ranges[blockIdx.z] in copy blockIdx.y (each BLAS keeps its pool range in every copy).
__global__ void k_oct_gather(uint32_t *__restrict__ oct, const uint32_t *__restrict__ nodes,
                             const uint32_t *__restrict__ perm, uint32_t n_total, const uint2 *__restrict__ ranges,
                             uint32_t n_nodes) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (ranges) {
        const uint2 r = ranges[blockIdx.z];
        if (r.x + i >= r.y) return;
        i += blockIdx.y * n_nodes + r.x;
    }
    if (i >= n_total) return;
    const uint32_t *src = nodes + (size_t)perm[i] * 8;
    uint32_t *dst = oct + (size_t)i * 8;
    dst[0] = src[0];
    dst[1] = src[1];
    dst[2] = src[2];
    dst[4] = src[3];
    dst[5] = src[4];
    dst[6] = src[5];
}

// One height level of a bottom-up BLAS refit (refitFlatBlasRange, js/gi_refit.js:34-165):
// order[0, n) lists nodes of equal height (leaves are height 0, an interior node is one
// above its taller child), so every child was finished by an earlier launch and the
// kernel boundary makes its bounds visible: no atomics and no fences. A leaf takes the
// bounds of its triangles (an empty leaf keeps its bounds, gi_refit.js:77-85); an interior
// node the union of its children (left = i + 1, right = the left child's miss).
__device__ __forceinline__ void refit_node(uint32_t *__restrict__ nodes, const uint32_t *__restrict__ tri_index,
                                           const float *__restrict__ vdata, uint32_t i) {
    float *nf = reinterpret_cast<float *>(nodes);
    const uint32_t info = nodes[(size_t)i * 8 + 7];
    float mn[3], mx[3];
    if (info != 0xFFFFFFFFu) {
        const uint32_t off = info & 0x00FFFFFFu, cnt = info >> 24;
        if (cnt == 0) return;
        for (int a = 0; a < 3; a++) {
            mn[a] = INFINITY;
            mx[a] = -INFINITY;
        }
        for (uint32_t t = off; t < off + cnt; t++) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float *p = vdata + (size_t)tri_index[(size_t)t * 3 + k] * 8;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    mn[a] = fminf(mn[a], p[a]);
                    mx[a] = fmaxf(mx[a], p[a]);
                }
            }
        }
    } else {
        const size_t l = (size_t)i + 1, r = nodes[l * 8 + 6];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = fminf(nf[l * 8 + a], nf[r * 8 + a]);
            mx[a] = fmaxf(nf[l * 8 + 3 + a], nf[r * 8 + 3 + a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        nf[(size_t)i * 8 + a] = mn[a];
        nf[(size_t)i * 8 + 3 + a] = mx[a];
    }
}

__global__ void k_refit_level(uint32_t *__restrict__ nodes, const uint32_t *__restrict__ tri_index,
                              const float *__restrict__ vdata, const uint32_t *__restrict__ order, uint32_t n) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    refit_node(nodes, tri_index, vdata, order[x]);
}

// One depth of a rebuilt tree: lv[2 d], lv[2 d + 1] are the depth's first slot in order
// and its node count (on the device, clamped to cap).
__global__ void k_refit_depth(uint32_t *__restrict__ nodes, const uint32_t *__restrict__ tri_index,
                              const float *__restrict__ vdata, const uint32_t *__restrict__ order,
                              const uint32_t *__restrict__ lv, uint32_t d, uint32_t cap) {
    const uint32_t off = min(lv[2 * d], cap), n = min(lv[2 * d + 1], cap - off);
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += gridDim.x * blockDim.x)
        refit_node(nodes, tri_index, vdata, order[off + x]);
}

}  // namespace

hipError_t launch_refit_levels(uint32_t *nodes, const uint32_t *tri_index, const float *vdata, const uint32_t *order,
                               const uint32_t *level_off, uint32_t levels, hipStream_t st) {
    for (uint32_t h = 0; h < levels; h++) {
        const uint32_t n = level_off[h + 1] - level_off[h];
        if (!n) continue;
        hipLaunchKernelGGL(k_refit_level, dim3((n + 255) / 256), dim3(256), 0, st, nodes, tri_index, vdata,
                           order + level_off[h], n);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_refit_depths(uint32_t *nodes, const uint32_t *tri_index, const float *vdata, const uint32_t *order,
                               const uint32_t *lv, uint32_t depths, uint32_t cap, const uint32_t *blocks,
                               hipStream_t st) {
    for (uint32_t d = depths; d-- > 0;) {
        hipLaunchKernelGGL(k_refit_depth, dim3(blocks[d]), dim3(256), 0, st, nodes, tri_index, vdata, order, lv, d,
                           cap);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_oct_gather(uint32_t *oct_nodes, const uint32_t *nodes, const uint32_t *perm, uint32_t n_total,
                             hipStream_t st) {
    if (!n_total) return hipSuccess;
    hipLaunchKernelGGL(k_oct_gather, dim3((n_total + 255) / 256), dim3(256), 0, st, oct_nodes, nodes, perm, n_total,
                       nullptr, 0u);
    return hipGetLastError();
}

hipError_t launch_oct_gather_ranges(uint32_t *oct_nodes, const uint32_t *nodes, const uint32_t *perm, uint32_t n_nodes,
                                    uint32_t copies, const uint2 *ranges, uint32_t n_ranges, uint32_t max_len,
                                    hipStream_t st) {
    if (!n_ranges || !max_len) return hipSuccess;
    hipLaunchKernelGGL(k_oct_gather, dim3((max_len + 255) / 256, copies, n_ranges), dim3(256), 0, st, oct_nodes, nodes,
                       perm, copies * n_nodes, ranges, n_nodes);
    return hipGetLastError();
}

hipError_t launch_pack_tris(const uint32_t *tri_index, const float *vdata, const uint32_t *tri_mat, const float *mats,
                            uint32_t tri_count, float4 *tri, hipStream_t st) {
    if (!tri_count) return hipSuccess;
    hipLaunchKernelGGL(k_pack_tris, dim3((tri_count + 255) / 256), dim3(256), 0, st, tri_index, vdata, tri_mat, mats,
                       tri_count, nullptr, tri);
    return hipGetLastError();
}

hipError_t launch_pack_tris_ranges(const uint32_t *tri_index, const float *vdata, const uint32_t *tri_mat,
                                   const float *mats, uint32_t tri_count, const uint2 *ranges, uint32_t n_ranges,
                                   uint32_t max_len, float4 *tri, hipStream_t st) {
    if (!n_ranges || !max_len) return hipSuccess;
    hipLaunchKernelGGL(k_pack_tris, dim3((max_len + 255) / 256, n_ranges), dim3(256), 0, st, tri_index, vdata, tri_mat,
                       mats, tri_count, ranges, tri);
    return hipGetLastError();
}

hipError_t launch_pack_tlas(const float *mats, uint32_t tlas_base, uint32_t tlas_count, uint4 *tlas, hipStream_t st) {
    if (!tlas_count) return hipSuccess;
    hipLaunchKernelGGL(k_pack_tlas, dim3((tlas_count + 255) / 256), dim3(256), 0, st, mats, tlas_base, tlas_count, tlas);
    return hipGetLastError();
}

hipError_t launch_pack_inst(const float *mats, uint32_t inst_base, uint32_t inst_count, float4 *inst, hipStream_t st) {
    if (!inst_count) return hipSuccess;
    hipLaunchKernelGGL(k_pack_inst, dim3((inst_count + 255) / 256), dim3(256), 0, st, mats, inst_base, inst_count,
                       inst);
    return hipGetLastError();
}

hipError_t launch_refit_tlas(float *mats, uint32_t tlas_base, uint32_t tlas_count, const float *inst_bounds,
                             const int32_t *parent, uint32_t *arrive, hipStream_t st) {
    if (!tlas_count) return hipSuccess;
    hipLaunchKernelGGL(k_refit_tlas, dim3((tlas_count + 255) / 256), dim3(256), 0, st, mats, tlas_base, tlas_count,
                       inst_bounds, parent, arrive);
    return hipGetLastError();
}
