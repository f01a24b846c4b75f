// k_dynamic.hip — per-frame scene edits that stay on the device.
//
// k_skin: linear-blend skinning of the registered vertices into the pooled vertexData
//   (positions and normals), from a rest pose, four joints and four weights per vertex
//   kept in HBM; a frame uploads only the bone palette. The reference receives skinned
//   positions from its host (CPU skinning) and re-gathers them into vertexData in
//   updateDeforms (js/spectral_scene.js:1831-1927); this produces the same buffer state
//   without the per-frame host gather and PCIe copy of every vertex.
// k_scatter_inst: instance records (inverse world rows, BLAS range, det sign, active),
//   their packed traversal copy and the double world matrices of the listed TLAS slots,
//   from one staged upload (updateTransforms, js/spectral_scene.js:1402-1514, rewrites
//   the same rows).
// k_fill_u32: clears arrival counters with a kernel, so the per-frame sequence holds no
//   memset node when it is captured into a graph.
#pragma clang fp contract(off)
#include "kernels.h"

namespace {

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// p' = sum_q w_q (B_q p), n' = normalize(sum_q w_q (B_q[3x3] n)), float, in this order
__global__ void k_skin(const uint32_t *__restrict__ vtx, const float4 *__restrict__ rest,
                       const uint2 *__restrict__ joints, const float4 *__restrict__ weights,
                       const float4 *__restrict__ bones, uint32_t n, float *__restrict__ vdata) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = rest[(size_t)i * 2], nr = rest[(size_t)i * 2 + 1];
    const uint2 jj = joints[i];
    const float4 ww = weights[i];
    const uint32_t j[4] = {jj.x & 0xFFFFu, jj.x >> 16, jj.y & 0xFFFFu, jj.y >> 16};
    const float w[4] = {ww.x, ww.y, ww.z, ww.w};
    float acc[3] = {0.f, 0.f, 0.f}, nacc[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float4 *B = bones + (size_t)j[q] * 3;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const float4 row = B[r];
            const float t = row.x * p.x + row.y * p.y + row.z * p.z + row.w;
            const float tn = row.x * nr.x + row.y * nr.y + row.z * nr.z;
            acc[r] = acc[r] + w[q] * t;
            nacc[r] = nacc[r] + w[q] * tn;
        }
    }
    const float len = sqrtf(nacc[0] * nacc[0] + nacc[1] * nacc[1] + nacc[2] * nacc[2]);
    float *o = vdata + (size_t)vtx[i] * 8;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        o[r] = acc[r];
        o[3 + r] = len > 0.f ? nacc[r] / len : nacc[r];
    }
}

__global__ void k_scatter_inst(const uint32_t *__restrict__ slots, const float *__restrict__ recs,
                               const double *__restrict__ worlds, uint32_t n, float *__restrict__ mats,
                               uint32_t inst_base, double *__restrict__ world, float4 *__restrict__ inst) {
    const uint32_t q = blockIdx.x;
    if (q >= n) return;
    const uint32_t s = slots[q], l = threadIdx.x;
    if (l < SBGI_MAT_STRIDE) mats[inst_base + (size_t)s * SBGI_MAT_STRIDE + l] = recs[(size_t)q * SBGI_MAT_STRIDE + l];
    if (l < 16) {
        world[(size_t)s * 16 + l] = worlds[(size_t)q * 16 + l];
        reinterpret_cast<float *>(inst)[(size_t)s * 16 + l] = recs[(size_t)q * SBGI_MAT_STRIDE + l];
    }
}

}  // namespace

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint32_t n, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_skin(const uint32_t *vtx, const float4 *rest, const uint2 *joints, const float4 *weights,
                       const float4 *bones, uint32_t n, float *vdata, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_skin, dim3((n + 255) / 256), dim3(256), 0, st, vtx, rest, joints, weights, bones, n, vdata);
    return hipGetLastError();
}

hipError_t launch_scatter_inst(const uint32_t *slots, const float *recs, const double *worlds, uint32_t n, float *mats,
                               uint32_t inst_base, double *world, float4 *inst, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_inst, dim3(n), dim3(64), 0, st, slots, recs, worlds, n, mats, inst_base, world, inst);
    return hipGetLastError();
}
