// k_build.hip — rebuild of BLASes on the device (sbgi_rebuild_blas).
//
// A refit keeps the tree of the rest pose; under large deformations its boxes swell. The
// rebuild replaces the tree of a BLAS node range in place, with the same node count, so
// instance records and the pool layout stay valid:
//   1. per triangle: AABB centroid; per BLAS: the centroids' bounds;
//   2. 30-bit Morton codes of the centroids, tagged with the BLAS index in the high bits,
//      sorted once for all BLASes (rocPRIM radix sort);
//   3. the triangles of each BLAS are permuted into Morton order; the range's
//      L = (N + 1) / 2 leaves take consecutive runs of them (T / L each, the first
//      T mod L one more);
//   4. the tree over the leaves is built top down, one launch per depth: a task covering
//      leaves [a, b) at node position p splits at the highest Morton bit in which its
//      first and last leaf differ (the middle when they are equal), and its children go
//      to p + 1 and p + 2 (m - a) (a subtree over k leaves has 2k - 1 nodes), so every
//      task of a depth is independent; the tasks, kept by depth, order the bottom-up
//      bounds pass (k_refit_depth, deepest first). Task counts stay on the device (a
//      depth table the launches read and extend), and the host launches a bound on the
//      depth (rebuild_depth_bound) instead of reading each depth's count back;
//   5. the walk copies: pool order (one copy), or the eight near-first octant copies,
//      placed top down from the subtree sizes and the child-box centres.
// No reference counterpart: the reference rebuilds on the CPU with three-mesh-bvh
// (js/spectral_scene.js:898-966) and keeps a refit otherwise (js/gi_refit.js).
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "kernels.h"

namespace {

// range descriptor: {pool root, node count N, first triangle t0, triangle count T}
__device__ __forceinline__ uint32_t find_range(const uint32_t *__restrict__ tri_prefix, uint32_t n_ranges,
                                               uint32_t g) {
    uint32_t lo = 0, hi = n_ranges;  // last range whose prefix <= g
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (tri_prefix[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t f2o(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) { return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o); }

// Centroids of every triangle of every range (g = global index over the ranges) and the
// per-range centroid bounds (6 order-mapped words per range, block-reduced, one atomic
// per block and word).
__global__ void k_centroids(const uint4 *__restrict__ ranges, const uint32_t *__restrict__ tri_prefix, uint32_t n_ranges,
                            uint32_t total, const uint32_t *__restrict__ tri_index, const float *__restrict__ vdata,
                            float4 *__restrict__ cen, uint32_t *__restrict__ cbounds) {
    __shared__ uint32_t s_b[6];
    __shared__ uint32_t s_r;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    // a block straddling two ranges reduces each lane into its own range with atomics
    const uint32_t r = g < total ? find_range(tri_prefix, n_ranges, g) : 0u;
    float c[3] = {0.f, 0.f, 0.f};
    if (g < total) {
        const uint32_t t = ranges[r].z + (g - tri_prefix[r]);
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 3; k++) {
            const float *p = vdata + (size_t)tri_index[(size_t)t * 3 + k] * 8;
            for (int a = 0; a < 3; a++) {
                mn[a] = fminf(mn[a], p[a]);
                mx[a] = fmaxf(mx[a], p[a]);
            }
        }
        for (int a = 0; a < 3; a++) c[a] = 0.5f * (mn[a] + mx[a]);
        cen[g] = make_float4(c[0], c[1], c[2], 0.f);
    }
    if (threadIdx.x == 0) {
        s_r = r;
        for (int q = 0; q < 3; q++) {
            s_b[q] = f2o(INFINITY);
            s_b[3 + q] = f2o(-INFINITY);
        }
    }
    __syncthreads();
    const bool mine = g < total && r == s_r;
    if (g < total && !mine) {  // the rare lanes of a second range in the block
        for (int q = 0; q < 3; q++) {
            atomicMin(&cbounds[r * 6 + q], f2o(c[q]));
            atomicMax(&cbounds[r * 6 + 3 + q], f2o(c[q]));
        }
    }
    if (mine)
        for (int q = 0; q < 3; q++) {
            atomicMin(&s_b[q], f2o(c[q]));
            atomicMax(&s_b[3 + q], f2o(c[q]));
        }
    __syncthreads();
    if (threadIdx.x < 6 && s_r < n_ranges) {
        if (threadIdx.x < 3) atomicMin(&cbounds[s_r * 6 + threadIdx.x], s_b[threadIdx.x]);
        else atomicMax(&cbounds[s_r * 6 + threadIdx.x], s_b[threadIdx.x]);
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// key = range << 30 | morton(centroid in the range's centroid box), value = local index
__global__ void k_morton(const uint32_t *__restrict__ tri_prefix, uint32_t n_ranges, uint32_t total,
                         const float4 *__restrict__ cen, const uint32_t *__restrict__ cbounds,
                         unsigned long long *__restrict__ keys, uint32_t *__restrict__ vals) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const uint32_t r = find_range(tri_prefix, n_ranges, g);
    const float4 c = cen[g];
    const float cc[3] = {c.x, c.y, c.z};
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        const float lo = o2f(cbounds[r * 6 + a]), hi = o2f(cbounds[r * 6 + 3 + a]);
        const float e = hi - lo;
        const float u = e > 0.f ? (cc[a] - lo) / e : 0.f;
        q[a] = (uint32_t)fminf(fmaxf(u * 1024.f, 0.f), 1023.f);
    }
    keys[g] = ((unsigned long long)r << 30) | (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[g] = g - tri_prefix[r];
}

// Triangle rows of every range into the sorted order (from a copy of the old rows).
__global__ void k_permute_tris(const uint4 *__restrict__ ranges, const uint32_t *__restrict__ tri_prefix,
                               uint32_t n_ranges, uint32_t total, const uint32_t *__restrict__ sorted_vals,
                               const uint32_t *__restrict__ old_index, const uint32_t *__restrict__ old_mat,
                               uint32_t *__restrict__ tri_index, uint32_t *__restrict__ tri_mat) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const uint32_t r = find_range(tri_prefix, n_ranges, g);
    const uint32_t base = tri_prefix[r], src = base + sorted_vals[g];  // position in the old copy
    const uint32_t dst = ranges[r].z + (g - base);
    for (int k = 0; k < 3; k++) tri_index[(size_t)dst * 3 + k] = old_index[(size_t)src * 3 + k];
    tri_mat[dst] = old_mat[src];
}

// leaf j of a range with T triangles in L leaves: [start, start + count)
__device__ __forceinline__ void leaf_run(uint32_t T, uint32_t L, uint32_t j, uint32_t &start, uint32_t &count) {
    const uint32_t base = T / L, rem = T % L;
    start = j * base + min(j, rem);
    count = base + (j < rem ? 1u : 0u);
}

// Depth table on the device: lv[2 d] is the first task of depth d in the task queue,
// lv[2 d + 1] its task count; lv[2 D + 2] is set when a child would fall past the queue.
// The launch for depth d writes lv[2 d + 2] and appends the tasks of depth d + 1 right
// after its own, so the queue index of a task is also its slot in `order`, and no count
// returns to the host between depths. Counts are clamped to the queue capacity.
__device__ __forceinline__ void depth_span(const uint32_t *lv, uint32_t d, uint32_t cap, uint32_t &off, uint32_t &n) {
    off = min(lv[2 * d], cap);
    n = min(lv[2 * d + 1], cap - off);
}

// One depth of the top-down build. Task = {node position, range, first leaf, end leaf}.
// codes: the sorted keys; a leaf's code is its first triangle's (30 low bits).
__global__ void k_build_level(const uint4 *__restrict__ ranges, const uint32_t *__restrict__ tri_prefix,
                              const unsigned long long *__restrict__ keys, uint4 *__restrict__ tasks, uint32_t cap,
                              uint32_t *__restrict__ lv, uint32_t d, uint32_t depths, uint32_t *__restrict__ nodes,
                              uint32_t *__restrict__ order) {
    uint32_t off, n;
    depth_span(lv, d, cap, off, n);
    const uint32_t noff = off + n;
    if (blockIdx.x == 0 && threadIdx.x == 0) lv[2 * d + 2] = noff;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint4 tk = tasks[off + i];
        const uint32_t pos = tk.x, r = tk.y, a = tk.z, b = tk.w;
        const uint4 rg = ranges[r];
        const uint32_t T = rg.w, L = (rg.y + 1) / 2;
        uint32_t *nd = nodes + (size_t)pos * 8;
        order[off + i] = pos;
        if (b - a == 1) {
            uint32_t st, cnt;
            leaf_run(T, L, a, st, cnt);
            nd[6] = pos + 1;
            nd[7] = (cnt << 24) | ((rg.z + st) & 0x00FFFFFFu);
            continue;
        }
        auto code = [&](uint32_t leaf) {
            uint32_t st, cnt;
            leaf_run(T, L, leaf, st, cnt);
            return (uint32_t)(keys[tri_prefix[r] + st] & 0x3FFFFFFFull);
        };
        const uint32_t ca = code(a), cb = code(b - 1);
        uint32_t m = a + (b - a) / 2;
        if (ca != cb) {
            // first leaf in (a, b) whose code has the highest differing bit set
            const uint32_t bit = 31u - __clz(ca ^ cb);
            uint32_t lo = a, hi = b - 1;  // code(lo) has the bit clear, code(hi) set
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if ((code(mid) >> bit) & 1u) hi = mid;
                else lo = mid;
            }
            m = hi;
        }
        nd[6] = pos + 2 * (b - a) - 1;
        nd[7] = 0xFFFFFFFFu;
        const uint32_t slot = noff + atomicAdd(&lv[2 * d + 3], 2u);
        if (d + 1 >= depths || slot + 2 > cap) {
            atomicOr(&lv[2 * depths + 2], 1u);
            continue;
        }
        tasks[slot] = make_uint4(pos + 1, r, a, m);
        tasks[slot + 1] = make_uint4(pos + 2 * (m - a), r, m, b);
    }
}

// Walk copies, top down by depth (the depth table's span of order lists the pool nodes of
// depth d): qpos[o][i] is pool node i's position in copy o. Near-first children as in the
// host layout: the child whose box centre projects lower on the octant direction.
__global__ void k_copy_positions(const uint32_t *__restrict__ nodes, const uint32_t *__restrict__ order,
                                 const uint32_t *__restrict__ lv, uint32_t d, uint32_t cap, uint32_t copies,
                                 uint32_t *__restrict__ qpos, uint32_t qstride) {
    uint32_t off, n;
    depth_span(lv, d, cap, off, n);
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += gridDim.x * blockDim.x) {
        const uint32_t i = order[off + x];
        const uint32_t *nd = nodes + (size_t)i * 8;
        if (nd[7] != 0xFFFFFFFFu) continue;
        const float *nf = reinterpret_cast<const float *>(nodes);
        const uint32_t l = i + 1, r = nodes[(size_t)l * 8 + 6];
        const uint32_t size_l = r - l;
        for (uint32_t o = 0; o < copies; o++) {
            const uint32_t q = qpos[(size_t)o * qstride + i];
            bool left_near = true;
            if (copies == 8) {
                const float dir[3] = {(o & 1) ? -1.f : 1.f, (o & 2) ? -1.f : 1.f, (o & 4) ? -1.f : 1.f};
                float cl = 0.f, cr = 0.f;
                for (int a = 0; a < 3; a++) {
                    cl += (nf[(size_t)l * 8 + a] + nf[(size_t)l * 8 + 3 + a]) * dir[a];
                    cr += (nf[(size_t)r * 8 + a] + nf[(size_t)r * 8 + 3 + a]) * dir[a];
                }
                left_near = cl <= cr;
            }
            const uint32_t size_r = nd[6] - r;
            qpos[(size_t)o * qstride + l] = left_near ? q + 1 : q + 1 + size_r;
            qpos[(size_t)o * qstride + r] = left_near ? q + 1 + size_l : q + 1;
        }
    }
}

// Each range's root keeps its pool position in every copy.
__global__ void k_copy_roots(const uint4 *__restrict__ ranges, uint32_t n_ranges, uint32_t copies,
                             uint32_t *__restrict__ qpos, uint32_t qstride) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_ranges) return;
    const uint32_t root = ranges[r].x;
    for (uint32_t o = 0; o < copies; o++) qpos[(size_t)o * qstride + root] = root;
}

// Write copy nodes {min.xyz, miss}, {max.xyz, payload} and perm for every pool node of the
// ranges (qpos indexed by pool node; miss in copy o = position + subtree size). n_built,
// when given, is the node count the build reached (the entries of order written).
__global__ void k_write_copies(const uint32_t *__restrict__ nodes, const uint32_t *__restrict__ order, uint32_t n,
                               const uint32_t *__restrict__ n_built, uint32_t copies,
                               const uint32_t *__restrict__ qpos, uint32_t qstride, uint32_t node_count,
                               uint32_t *__restrict__ walk, uint32_t *__restrict__ perm) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (n_built ? min(n, *n_built) : n)) return;
    const uint32_t i = order[x];
    const uint32_t *src = nodes + (size_t)i * 8;
    const uint32_t size = src[6] - i;
    for (uint32_t o = 0; o < copies; o++) {
        const uint32_t q = copies == 1 ? i : qpos[(size_t)o * qstride + i];
        uint32_t *dst = walk + ((size_t)o * node_count + q) * 8;
        dst[0] = src[0];
        dst[1] = src[1];
        dst[2] = src[2];
        dst[3] = q + size;
        dst[4] = src[3];
        dst[5] = src[4];
        dst[6] = src[5];
        dst[7] = src[7];
        perm[(size_t)o * node_count + q] = i;
    }
}

}  // namespace

size_t rebuild_sort_bytes(uint32_t total) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (unsigned long long *)nullptr, (unsigned long long *)nullptr,
                                    (uint32_t *)nullptr, (uint32_t *)nullptr, total, 0u, 64u);
    return bytes;
}

hipError_t launch_rebuild_prepare(const uint4 *ranges, const uint32_t *tri_prefix, uint32_t n_ranges, uint32_t total,
                                  const uint32_t *tri_index, const float *vdata, float4 *cen, uint32_t *cbounds,
                                  unsigned long long *keys, uint32_t *vals, unsigned long long *keys_sorted,
                                  uint32_t *vals_sorted, void *sort_tmp, size_t sort_bytes, uint32_t end_bit,
                                  hipStream_t st) {
    if (!total) return hipSuccess;
    const uint32_t blocks = (total + 255) / 256;
    hipLaunchKernelGGL(k_centroids, dim3(blocks), dim3(256), 0, st, ranges, tri_prefix, n_ranges, total, tri_index,
                       vdata, cen, cbounds);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_morton, dim3(blocks), dim3(256), 0, st, tri_prefix, n_ranges, total, cen, cbounds, keys, vals);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys, keys_sorted, vals, vals_sorted, total, 0u, end_bit, st);
}

hipError_t launch_permute_tris(const uint4 *ranges, const uint32_t *tri_prefix, uint32_t n_ranges, uint32_t total,
                               const uint32_t *sorted_vals, const uint32_t *old_index, const uint32_t *old_mat,
                               uint32_t *tri_index, uint32_t *tri_mat, hipStream_t st) {
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_permute_tris, dim3((total + 255) / 256), dim3(256), 0, st, ranges, tri_prefix, n_ranges, total,
                       sorted_vals, old_index, old_mat, tri_index, tri_mat);
    return hipGetLastError();
}

uint32_t rebuild_depth_bound(uint32_t max_leaves) {
    // a radix split lowers the highest differing Morton bit (30 bits), the middle splits
    // of equal codes halve the leaf count, and a chain over L leaves is L deep
    uint32_t k = 0;
    while (k < 32 && (1ull << k) < max_leaves) k++;
    return std::max(1u, std::min(31u + k, max_leaves));
}

hipError_t launch_build_depths(const uint4 *ranges, const uint32_t *tri_prefix, const unsigned long long *keys,
                               uint4 *tasks, uint32_t cap, uint32_t *lv, uint32_t depths, const uint32_t *blocks,
                               uint32_t *nodes, uint32_t *order, hipStream_t st) {
    for (uint32_t d = 0; d < depths; d++) {
        hipLaunchKernelGGL(k_build_level, dim3(blocks[d]), dim3(256), 0, st, ranges, tri_prefix, keys, tasks, cap, lv,
                           d, depths, nodes, order);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_copy_roots(const uint4 *ranges, uint32_t n_ranges, uint32_t copies, uint32_t *qpos, uint32_t qstride,
                             hipStream_t st) {
    if (!n_ranges) return hipSuccess;
    hipLaunchKernelGGL(k_copy_roots, dim3((n_ranges + 255) / 256), dim3(256), 0, st, ranges, n_ranges, copies, qpos,
                       qstride);
    return hipGetLastError();
}

hipError_t launch_copy_positions(const uint32_t *nodes, const uint32_t *order, const uint32_t *lv, uint32_t depths,
                                 uint32_t cap, const uint32_t *blocks, uint32_t copies, uint32_t *qpos,
                                 uint32_t qstride, hipStream_t st) {
    for (uint32_t d = 0; d < depths; d++) {
        hipLaunchKernelGGL(k_copy_positions, dim3(blocks[d]), dim3(256), 0, st, nodes, order, lv, d, cap, copies, qpos,
                           qstride);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_write_copies(const uint32_t *nodes, const uint32_t *order, uint32_t n, const uint32_t *n_built,
                               uint32_t copies, const uint32_t *qpos, uint32_t qstride, uint32_t node_count,
                               uint32_t *walk, uint32_t *perm, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_write_copies, dim3((n + 255) / 256), dim3(256), 0, st, nodes, order, n, n_built, copies, qpos,
                       qstride, node_count, walk, perm);
    return hipGetLastError();
}
