// k_tlas.hip — the per-frame TLAS on the device.
//
// k_inst_bounds: world AABB of every TLAS slot, Box3.applyMatrix4 in double of the
//   instance's BLAS root bounds (js/spectral_scene.js:1301-1306); an inactive slot is the
//   point at its translation. It also writes the bounds inflated by a relative 1e-5 for
//   the SAH build, so that rounding of a world matrix can never make a box smaller than
//   its instance.
// k_build_tlas: binned-SAH TLAS (32 bins per axis, one slot per leaf) over the inflated
//   bounds, written in the walk's node format {min.xyz, miss}, {max.xyz, payload}. The
//   reference keeps a median-split TLAS built once on the CPU and refits it
//   (js/spectral_scene.js:973-1003); rebuilding with SAH every frame keeps the tree
//   good as instances move.
//
// The build is one workgroup of 16 waves. Node positions are depth-first (left child at
//   pos + 1, right child at pos + 2·|left items|, miss at pos + 2·|items| − 1: a subtree
//   over m slots has 2m − 1 nodes), so every node of a level is independent and a level
//   is processed in parallel, one wave per node, with the slot ids partitioned stably
//   from one ping-pong array into the other. Bin counts and bounds are LDS atomics (min
//   and max through an order-preserving float→u32 map); the SAH sweep is a segmented
//   prefix/suffix scan across lanes with costs in double. Every decision (bin index, sweep order, strict "<" across axes and
//   bins, stable partition, middle split when no axis has extent) is deterministic, so
//   the tree does not depend on the schedule of the waves.
#pragma clang fp contract(off)
#include "kernels.h"

namespace {

constexpr int kTlasBins = 32;
constexpr int kTlasBlock = 1024;
constexpr int kTlasWaves = kTlasBlock / 64;

__global__ void k_inst_bounds(const float *__restrict__ mats, uint32_t inst_base, uint32_t n,
                              const uint32_t *__restrict__ nodes, const double *__restrict__ world,
                              float *__restrict__ bounds, float *__restrict__ padded) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *rec = mats + inst_base + (size_t)i * SBGI_MAT_STRIDE;
    const double *e = world + (size_t)i * 16;
    float out[6];
    if (!(rec[15] > 0.5f)) {
        for (int a = 0; a < 3; a++) out[a] = out[3 + a] = (float)e[12 + a];
    } else {
        const float *rb = reinterpret_cast<const float *>(nodes + (size_t)(uint32_t)rec[12] * 8);
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c = 0; c < 8; c++) {
            const double x = (c & 1) ? rb[3] : rb[0];
            const double y = (c & 2) ? rb[4] : rb[1];
            const double z = (c & 4) ? rb[5] : rb[2];
            const double w = 1 / (e[3] * x + e[7] * y + e[11] * z + e[15]);
            const double p[3] = {(e[0] * x + e[4] * y + e[8] * z + e[12]) * w,
                                 (e[1] * x + e[5] * y + e[9] * z + e[13]) * w,
                                 (e[2] * x + e[6] * y + e[10] * z + e[14]) * w};
            for (int a = 0; a < 3; a++) {
                mn[a] = p[a] < mn[a] ? p[a] : mn[a];
                mx[a] = mx[a] < p[a] ? p[a] : mx[a];
            }
        }
        for (int a = 0; a < 3; a++) {
            out[a] = (float)mn[a];
            out[3 + a] = (float)mx[a];
        }
    }
    for (int a = 0; a < 6; a++) bounds[(size_t)i * 6 + a] = out[a];
    for (int a = 0; a < 3; a++) {
        const float pad = 1e-5f * (fabsf(out[3 + a] - out[a]) + fabsf(out[a]) + fabsf(out[3 + a])) + 1e-6f;
        padded[(size_t)i * 6 + a] = out[a] - pad;
        padded[(size_t)i * 6 + 3 + a] = out[3 + a] + pad;
    }
}

// order-preserving map of finite floats and infinities onto u32 (for LDS atomic min/max)
__device__ __forceinline__ uint32_t f2o(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

__device__ __forceinline__ float fmin_keep(float acc, float v) { return v < acc ? v : acc; }
__device__ __forceinline__ float fmax_keep(float acc, float v) { return acc < v ? v : acc; }

// surface area of a box in double; an empty or inverted box has area 0
__device__ __forceinline__ double box_area(const float *mn, const float *mx) {
    double d[3];
    for (int a = 0; a < 3; a++) {
        d[a] = (double)mx[a] - (double)mn[a];
        if (!(d[a] >= 0)) return 0.0;
    }
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}

__device__ __forceinline__ int bin_of(float c, float cmn, float ext) {
    int k = (int)((c - cmn) / ext * kTlasBins);
    return min(max(k, 0), kTlasBins - 1);
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// bins of one wave: per axis and bin {count, min.xyz, max.xyz} (min/max as f2o words)
constexpr int kBinWords = 7;

// Up to kTlasLdsSlots slots, the boxes and both id arrays live in LDS for the whole build
// (32 bytes a slot in dynamic LDS): every level reads each box three times (bounds,
// binning, partition) and the ids twice, so the chain of dependent loads per task runs at
// LDS latency instead of L2 latency. Larger TLASes read them from global memory.
constexpr uint32_t kTlasLdsSlots = 2048;

template <bool LDS>
__global__ __launch_bounds__(kTlasBlock) void k_build_tlas(const float *__restrict__ bnd_g, uint32_t n,
                                                           uint32_t *__restrict__ ids0_g, uint32_t *__restrict__ ids1_g,
                                                           uint4 *__restrict__ task0, uint4 *__restrict__ task1,
                                                           uint4 *__restrict__ tlas) {
    __shared__ uint32_t s_count[2];
    __shared__ uint32_t s_bins[kTlasWaves][3 * kTlasBins * kBinWords];
    extern __shared__ __align__(16) uint32_t s_slots[];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const float *bnd;
    uint32_t *ids0, *ids1;
    if constexpr (LDS) {
        float *sb = reinterpret_cast<float *>(s_slots);
        for (uint32_t i = tid; i < n * 6; i += kTlasBlock) sb[i] = bnd_g[i];
        bnd = sb;
        ids0 = s_slots + (size_t)n * 6;
        ids1 = ids0 + n;
    } else {
        bnd = bnd_g;
        ids0 = ids0_g;
        ids1 = ids1_g;
    }
    for (uint32_t i = tid; i < n; i += kTlasBlock) ids0[i] = i;
    if (tid == 0) {
        task0[0] = make_uint4(0u, 0u, n, 0u);
        s_count[0] = 1u;
        s_count[1] = 0u;
    }
    __syncthreads();
    uint32_t *bins = s_bins[wave];
    for (uint32_t level = 0;; level++) {
        const uint32_t cur = level & 1u, nxt = cur ^ 1u;
        const uint32_t ntask = s_count[cur];
        __syncthreads();  // every thread has read the task count; s_count[nxt] is zero
        if (ntask == 0) break;
        const uint4 *tasks = cur ? task1 : task0;
        uint4 *next = cur ? task0 : task1;
        const uint32_t *src = cur ? ids1 : ids0;
        uint32_t *dst = cur ? ids0 : ids1;
        for (uint32_t t = wave; t < ntask; t += kTlasWaves) {
            const uint4 tk = tasks[t];
            const uint32_t pos = tk.x, lo = tk.y, hi = tk.z, m = hi - lo;
            float bmn[3] = {INFINITY, INFINITY, INFINITY}, bmx[3] = {-INFINITY, -INFINITY, -INFINITY};
            float cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t i = lo + lane; i < hi; i += 64) {
                const float *bb = bnd + (size_t)src[i] * 6;
                for (int a = 0; a < 3; a++) {
                    bmn[a] = fmin_keep(bmn[a], bb[a]);
                    bmx[a] = fmax_keep(bmx[a], bb[3 + a]);
                    const float c = 0.5f * (bb[a] + bb[3 + a]);
                    cmn[a] = fmin_keep(cmn[a], c);
                    cmx[a] = fmax_keep(cmx[a], c);
                }
            }
            for (int a = 0; a < 3; a++) {
                bmn[a] = wave_min(bmn[a]);
                bmx[a] = wave_max(bmx[a]);
                cmn[a] = wave_min(cmn[a]);
                cmx[a] = wave_max(cmx[a]);
            }
            if (m == 1) {
                if (lane == 0) {
                    tlas[(size_t)pos * 2] =
                        make_uint4(__float_as_uint(bmn[0]), __float_as_uint(bmn[1]), __float_as_uint(bmn[2]), pos + 1);
                    tlas[(size_t)pos * 2 + 1] = make_uint4(__float_as_uint(bmx[0]), __float_as_uint(bmx[1]),
                                                           __float_as_uint(bmx[2]), (1u << 24) | (src[lo] & 0x00FFFFFFu));
                }
                continue;
            }
            const float ext[3] = {cmx[0] - cmn[0], cmx[1] - cmn[1], cmx[2] - cmn[2]};
            for (uint32_t j = lane; j < 3 * kTlasBins; j += 64) {
                uint32_t *b = bins + j * kBinWords;
                b[0] = 0u;
                b[1] = b[2] = b[3] = f2o(INFINITY);
                b[4] = b[5] = b[6] = f2o(-INFINITY);
            }
            wave_sync();
            for (uint32_t i = lo + lane; i < hi; i += 64) {
                const float *bb = bnd + (size_t)src[i] * 6;
                for (int a = 0; a < 3; a++) {
                    if (!(ext[a] > 0)) continue;
                    const int k = bin_of(0.5f * (bb[a] + bb[3 + a]), cmn[a], ext[a]);
                    uint32_t *b = bins + (a * kTlasBins + k) * kBinWords;
                    atomicAdd(b, 1u);
                    for (int q = 0; q < 3; q++) {
                        atomicMin(b + 1 + q, f2o(bb[q]));
                        atomicMax(b + 4 + q, f2o(bb[3 + q]));
                    }
                }
            }
            wave_sync();
            // SAH sweep across lanes: lane k of a 32-lane segment owns bin k of one axis
            // (axes 0 and 1, then axis 2), an inclusive prefix scan gives the left side
            // union(bins[0..k]) and a suffix scan the right side union(bins[k+1..31]); the
            // cost of the split after bin k is area(L)|L| + area(R)|R|, and the wave arg-min
            // over (cost, 32 axis + k) keeps the first minimum in axis, then bin, order.
            double best = INFINITY;
            int best_idx = 0x7FFFFFFF;
            const int k = (int)(lane & 31u);
            for (int pass = 0; pass < 2; pass++) {
                const int a = pass == 0 ? (int)(lane >> 5) : 2;
                const float e = a == 0 ? ext[0] : (a == 1 ? ext[1] : ext[2]);
                const bool axis_ok = (pass == 0 || lane < 32) && e > 0;
                uint32_t pc = 0;
                float pmn[3] = {INFINITY, INFINITY, INFINITY}, pmx[3] = {-INFINITY, -INFINITY, -INFINITY};
                if (axis_ok) {
                    const uint32_t *b = bins + (a * kTlasBins + k) * kBinWords;
                    pc = b[0];
                    for (int q = 0; q < 3; q++) {
                        pmn[q] = o2f(b[1 + q]);
                        pmx[q] = o2f(b[4 + q]);
                    }
                }
                uint32_t sc = pc;
                float smn[3] = {pmn[0], pmn[1], pmn[2]}, smx[3] = {pmx[0], pmx[1], pmx[2]};
                for (int d = 1; d < kTlasBins; d <<= 1) {
                    const uint32_t oc = (uint32_t)__shfl_up((int)pc, d, kTlasBins);
                    const uint32_t uc = (uint32_t)__shfl_down((int)sc, d, kTlasBins);
                    float omn[3], omx[3], umn[3], umx[3];
                    for (int q = 0; q < 3; q++) {
                        omn[q] = __shfl_up(pmn[q], d, kTlasBins);
                        omx[q] = __shfl_up(pmx[q], d, kTlasBins);
                        umn[q] = __shfl_down(smn[q], d, kTlasBins);
                        umx[q] = __shfl_down(smx[q], d, kTlasBins);
                    }
                    if (k >= d) {
                        pc += oc;
                        for (int q = 0; q < 3; q++) {
                            pmn[q] = fmin_keep(pmn[q], omn[q]);
                            pmx[q] = fmax_keep(pmx[q], omx[q]);
                        }
                    }
                    if (k + d < kTlasBins) {
                        sc += uc;
                        for (int q = 0; q < 3; q++) {
                            smn[q] = fmin_keep(smn[q], umn[q]);
                            smx[q] = fmax_keep(smx[q], umx[q]);
                        }
                    }
                }
                // right side of the split after bin k = the suffix starting at bin k + 1
                const uint32_t rc = (uint32_t)__shfl_down((int)sc, 1, kTlasBins);
                float rmn[3], rmx[3];
                for (int q = 0; q < 3; q++) {
                    rmn[q] = __shfl_down(smn[q], 1, kTlasBins);
                    rmx[q] = __shfl_down(smx[q], 1, kTlasBins);
                }
                if (axis_ok && k < kTlasBins - 1 && pc && rc) {
                    const double cost = box_area(pmn, pmx) * pc + box_area(rmn, rmx) * rc;
                    const int idx = a * kTlasBins + k;
                    if (cost < best || (cost == best && idx < best_idx)) {
                        best = cost;
                        best_idx = idx;
                    }
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double oc = __shfl_xor(best, o);
                const int oi = __shfl_xor(best_idx, o);
                if (oc < best || (oc == best && oi < best_idx)) {
                    best = oc;
                    best_idx = oi;
                }
            }
            const int best_axis = best_idx == 0x7FFFFFFF ? -1 : best_idx / kTlasBins;
            const int bbin = best_idx == 0x7FFFFFFF ? -1 : best_idx % kTlasBins;
            uint32_t mid;
            if (best_axis < 0) {
                mid = lo + m / 2;
                for (uint32_t i = lo + lane; i < hi; i += 64) dst[i] = src[i];
            } else {
                const float e = best_axis == 0 ? ext[0] : (best_axis == 1 ? ext[1] : ext[2]);
                const float cm = best_axis == 0 ? cmn[0] : (best_axis == 1 ? cmn[1] : cmn[2]);
                auto goes_left = [&](uint32_t id) {
                    const float *bb = bnd + (size_t)id * 6;
                    return bin_of(0.5f * (bb[best_axis] + bb[3 + best_axis]), cm, e) <= bbin;
                };
                // count with the same predicate that scatters, so every write stays in [lo, hi)
                uint32_t nleft = 0;
                for (uint32_t base = lo; base < hi; base += 64) {
                    const uint32_t i = base + lane;
                    nleft += __popcll(__ballot(i < hi && goes_left(src[i])));
                }
                uint32_t lbase = 0, rbase = 0;
                const uint64_t below = (1ull << lane) - 1ull;
                for (uint32_t base = lo; base < hi; base += 64) {
                    const uint32_t i = base + lane;
                    const bool valid = i < hi;
                    const uint32_t id = valid ? src[i] : 0u;
                    const bool left = valid && goes_left(id);
                    const uint64_t bl = __ballot(left), bv = __ballot(valid), br = bv & ~bl;
                    if (valid) {
                        if (left)
                            dst[lo + lbase + __popcll(bl & below)] = id;
                        else
                            dst[lo + nleft + rbase + __popcll(br & below)] = id;
                    }
                    lbase += __popcll(bl);
                    rbase += __popcll(br);
                }
                mid = lo + nleft;
                if (mid == lo || mid == hi) mid = lo + m / 2;
            }
            if (lane == 0) {
                tlas[(size_t)pos * 2] = make_uint4(__float_as_uint(bmn[0]), __float_as_uint(bmn[1]),
                                                   __float_as_uint(bmn[2]), pos + 2 * m - 1);
                tlas[(size_t)pos * 2 + 1] =
                    make_uint4(__float_as_uint(bmx[0]), __float_as_uint(bmx[1]), __float_as_uint(bmx[2]), 0xFFFFFFFFu);
                const uint32_t slot = atomicAdd(&s_count[nxt], 2u);
                next[slot] = make_uint4(pos + 1, lo, mid, 0u);
                next[slot + 1] = make_uint4(pos + 2 * (mid - lo), mid, hi, 0u);
            }
        }
        __syncthreads();  // this level's nodes, partitions and appended tasks are complete
        if (tid == 0) s_count[cur] = 0u;
    }
}

// Near-first copies of the SAH TLAS, one per ray direction octant o (sign bits of x, y, z):
// in copy o every interior node's children are ordered by the projection of their box
// centres on the octant direction, nearer first, so a closest-hit walk meets near instances
// first and culls far ones against its shrinking t. Positions follow the pre-order of the
// source (a node's position is known before its children's): thread o walks the nodes in
// index order in LDS, then every thread writes copy nodes {min.xyz, position + subtree size},
// {max.xyz, payload}. Only TLASes of at most kTlasOctMax nodes get copies.
__global__ __launch_bounds__(256) void k_tlas_octants(const uint4 *__restrict__ tlas, uint32_t m, uint4 *__restrict__ out) {
    __shared__ uint4 s_nodes[2 * kTlasOctMax];
    __shared__ uint16_t s_pos[8][kTlasOctMax];
    for (uint32_t i = threadIdx.x; i < 2 * m; i += blockDim.x) s_nodes[i] = tlas[i];
    for (uint32_t i = threadIdx.x; i < 8 * kTlasOctMax; i += blockDim.x) s_pos[i / kTlasOctMax][i % kTlasOctMax] = 0xFFFFu;
    __syncthreads();
    if (threadIdx.x < 8) {
        const uint32_t o = threadIdx.x;
        const float sx = (o & 1) ? -1.f : 1.f, sy = (o & 2) ? -1.f : 1.f, sz = (o & 4) ? -1.f : 1.f;
        uint16_t *q = s_pos[o];
        q[0] = 0;
        for (uint32_t i = 0; i < m; i++) {
            const uint4 a = s_nodes[2 * i], b = s_nodes[2 * i + 1];
            if (b.w != 0xFFFFFFFFu) continue;
            const uint32_t l = i + 1, r = l < m ? s_nodes[2 * l].w : 0u;  // left child's miss = right child
            if (!(r > l && r < m)) continue;  // not a valid threaded layout: left as placed
            const uint4 la = s_nodes[2 * l], lb = s_nodes[2 * l + 1], ra = s_nodes[2 * r], rb = s_nodes[2 * r + 1];
            const float cl = (__uint_as_float(la.x) + __uint_as_float(lb.x)) * sx +
                             (__uint_as_float(la.y) + __uint_as_float(lb.y)) * sy +
                             (__uint_as_float(la.z) + __uint_as_float(lb.z)) * sz;
            const float cr = (__uint_as_float(ra.x) + __uint_as_float(rb.x)) * sx +
                             (__uint_as_float(ra.y) + __uint_as_float(rb.y)) * sy +
                             (__uint_as_float(ra.z) + __uint_as_float(rb.z)) * sz;
            const uint32_t size_l = r - l, size_r = a.w - r;
            const bool left_near = cl <= cr;
            q[l] = (uint16_t)(left_near ? q[i] + 1u : q[i] + 1u + size_r);
            q[r] = (uint16_t)(left_near ? q[i] + 1u + size_l : q[i] + 1u);
        }
    }
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < 8 * m; x += blockDim.x) {
        const uint32_t o = x / m, i = x - o * m;
        const uint32_t q = s_pos[o][i];
        if (q >= m) continue;  // unreached (only in an invalid layout)
        const uint4 a = s_nodes[2 * i], b = s_nodes[2 * i + 1];
        uint4 *d = out + 2 * ((size_t)o * m + q);
        d[0] = make_uint4(a.x, a.y, a.z, q + (a.w - i));
        d[1] = b;
    }
}

}  // namespace

hipError_t launch_tlas_octants(const uint4 *tlas, uint32_t m, uint4 *out, hipStream_t st) {
    if (!m || m > kTlasOctMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_tlas_octants, dim3(1), dim3(256), 0, st, tlas, m, out);
    return hipGetLastError();
}

size_t tlas_build_scratch_bytes(uint32_t n) { return (size_t)n * (2 * sizeof(uint32_t) + 2 * sizeof(uint4)); }

hipError_t launch_inst_bounds(const float *mats, uint32_t inst_base, uint32_t n, const uint32_t *nodes,
                              const double *world, float *bounds, float *padded, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_inst_bounds, dim3((n + 255) / 256), dim3(256), 0, st, mats, inst_base, n, nodes, world, bounds,
                       padded);
    return hipGetLastError();
}

hipError_t launch_build_tlas(const float *padded, uint32_t n, void *scratch, uint4 *tlas, hipStream_t st) {
    if (!n) return hipSuccess;
    // tasks first: scratch is 256-byte aligned and the uint4 lists need 16
    uint4 *task0 = static_cast<uint4 *>(scratch), *task1 = task0 + n;
    uint32_t *ids0 = reinterpret_cast<uint32_t *>(task1 + n), *ids1 = ids0 + n;
    if (n <= kTlasLdsSlots)
        hipLaunchKernelGGL(k_build_tlas<true>, dim3(1), dim3(kTlasBlock), (size_t)n * 32, st, padded, n, ids0, ids1,
                           task0, task1, tlas);
    else
        hipLaunchKernelGGL(k_build_tlas<false>, dim3(1), dim3(kTlasBlock), 0, st, padded, n, ids0, ids1, task0, task1,
                           tlas);
    return hipGetLastError();
}
