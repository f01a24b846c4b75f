// ceilings.hip — measurement tool (not part of the library): the ceilings of the walk
// kernels' rooflines on this GPU, each from a kernel that streams one kind of work at full
// occupancy (4,096 blocks of 256 threads) so that the work's issue or transfer rate is its
// only limit.
//
// VALU issue, one instruction class per kernel, 8 independent chains per lane:
//   0 v_fma_f32    1 v_min_f32 / v_max_f32    2 v_add_u32 + shift + xor
//   3 v_cmp_f32 + v_cndmask    4 v_rcp_f32 (+ v_add_f32)    5 v_mul_f32    6 v_add_f32
//   7 v_cvt_f32_u32 (+ xor)    8 64-bit integer add / shift    9 xor + shift
// `rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_VALU_<class> ...` over the same run counts each
// kernel's wave-instructions by class; count / best time is the class's issue rate.
//
// Cache bandwidth, 16-byte loads, four in flight per lane:
//   l2: every block reads a 2 MiB buffer (resident in each XCD's 4 MiB L2, 64 times the
//       32 KiB L1) from its own starting offset, so the loads miss L1 and hit L2;
//   l1: every block re-reads its own 2 KiB slice of the buffer (the blocks of one CU
//       together stay within its L1), so the loads hit L1.
// Bytes moved / best time is the ceiling; TCP/TCC counters over the same run give the
// bytes per counted request.
//
// Every store is out[block * 256 + thread]; every load index is masked into the buffer.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {
constexpr int kBlocks = 4096, kThreads = 256;
constexpr uint32_t kBufVec = (2u << 20) / 16;  // float4 elements in the 2 MiB buffer
constexpr uint32_t kL1Vec = (2u << 10) / 16;   // float4 elements in a block's L1 slice

template <int KIND>
__global__ void __launch_bounds__(kThreads) k_stream(float *out, float a, float b, int iters) {
    float x[8];
    uint32_t u[8];
    uint64_t q[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        x[j] = threadIdx.x * 1e-3f + j;
        u[j] = threadIdx.x + j * 77u;
        q[j] = ((uint64_t)threadIdx.x << 33) + (uint64_t)j * 0x9E3779B97F4A7C15ull;
    }
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int r = 0; r < 8; r++) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (KIND == 0) x[j] = fmaf(x[j], a, b);
                if (KIND == 1) x[j] = fminf(fmaxf(x[j], x[(j + 1) & 7]), x[(j + 3) & 7]);
                if (KIND == 2) u[j] = (u[j] + u[(j + 1) & 7]) ^ (u[(j + 2) & 7] >> 1);
                if (KIND == 3) x[j] = x[(j + 1) & 7] < x[j] ? x[j] : x[(j + 5) & 7];
                if (KIND == 4) x[j] = __builtin_amdgcn_rcpf(x[j] + a);
                if (KIND == 5) x[j] = x[j] * a;
                if (KIND == 6) x[j] = x[j] + x[(j + 3) & 7];
                if (KIND == 7) u[j] = __float_as_uint((float)u[j]) ^ u[(j + 1) & 7];
                if (KIND == 8) q[j] = (q[j] + q[(j + 1) & 7]) << (j + 1);
                if (KIND == 9) u[j] = u[j] ^ (u[(j + 3) & 7] >> (j + 1));
            }
        }
    }
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < 8; j++) acc += x[j] + (float)u[j] + (float)(uint32_t)(q[j] >> 32);
    out[blockIdx.x * kThreads + threadIdx.x] = acc;
}

// L2 (L1 = false) or L1 (L1 = true) read stream over buf (kBufVec float4 elements).
template <bool L1>
__global__ void __launch_bounds__(kThreads) k_cache_stream(const float4 *__restrict__ buf, float *out, int iters) {
    const uint32_t t = threadIdx.x;
    // each block starts at its own 64 KiB-aligned offset (L2) or owns a 2 KiB slice (L1)
    const uint32_t base = L1 ? (blockIdx.x * kL1Vec) & (kBufVec - 1) : (blockIdx.x * 4096u) & (kBufVec - 1);
    const uint32_t span = L1 ? kL1Vec : kBufVec;
    uint32_t acc = 0;
    for (int i = 0; i < iters; i++) {
        // L2: consecutive 16 KiB steps through the buffer; L1: a rotating lane offset inside
        // the slice (an index that changes every iteration, so no load is loop-invariant)
        const uint32_t o = L1 ? (uint32_t)i * 37u : (uint32_t)i * 4u * kThreads;
        float4 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t j = L1 ? o + (uint32_t)k * 29u + t : o + (uint32_t)k * kThreads + t;
            v[k] = buf[(base + (j & (span - 1))) & (kBufVec - 1)];
        }
#pragma unroll
        for (int k = 0; k < 4; k++)
            acc ^= __float_as_uint(v[k].x) ^ __float_as_uint(v[k].y) ^ __float_as_uint(v[k].z) ^
                   __float_as_uint(v[k].w);
    }
    out[blockIdx.x * kThreads + t] = (float)acc;
}
}  // namespace

int main() {
    float *out = nullptr;
    float4 *buf = nullptr;
    if (hipMalloc(&out, (size_t)kBlocks * kThreads * sizeof(float)) != hipSuccess) return 1;
    if (hipMalloc(&buf, (size_t)kBufVec * sizeof(float4)) != hipSuccess) return 1;
    if (hipMemset(buf, 0x3c, (size_t)kBufVec * sizeof(float4)) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    auto best_of_3 = [&](auto launch, float &best) -> int {
        best = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            if (hipEventRecord(e0) != hipSuccess) return 1;
            launch();
            if (hipGetLastError() != hipSuccess || hipEventRecord(e1) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess)
                return 1;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
            best = ms < best ? ms : best;
        }
        return 0;
    };
    const char *names[10] = {"fma_f32", "min_max_f32", "int32_add_shift_xor", "cmp_cndmask", "rcp_f32",
                             "mul_f32", "add_f32",     "cvt",                 "int64",       "xor_shift"};
    const int iters = 256;
    int rc = 0;
    auto valu = [&](auto kern, int kind) {
        float ms;
        rc |= best_of_3([&] { hipLaunchKernelGGL(kern, dim3(kBlocks), dim3(kThreads), 0, 0, out, 0.999f, 1e-3f, iters); },
                        ms);
        printf("valu %d %s %.4f ms\n", kind, names[kind], ms);
    };
    valu(k_stream<0>, 0);
    valu(k_stream<1>, 1);
    valu(k_stream<2>, 2);
    valu(k_stream<3>, 3);
    valu(k_stream<4>, 4);
    valu(k_stream<5>, 5);
    valu(k_stream<6>, 6);
    valu(k_stream<7>, 7);
    valu(k_stream<8>, 8);
    valu(k_stream<9>, 9);
    const int citers = 512;
    const double bytes = (double)kBlocks * kThreads * citers * 4 * sizeof(float4);
    float ms;
    rc |= best_of_3([&] { hipLaunchKernelGGL(k_cache_stream<false>, dim3(kBlocks), dim3(kThreads), 0, 0, buf, out, citers); },
                    ms);
    printf("cache l2 %.0f bytes %.4f ms %.1f GB/s\n", bytes, ms, bytes / (ms * 1e-3) / 1e9);
    rc |= best_of_3([&] { hipLaunchKernelGGL(k_cache_stream<true>, dim3(kBlocks), dim3(kThreads), 0, 0, buf, out, citers); },
                    ms);
    printf("cache l1 %.0f bytes %.4f ms %.1f GB/s\n", bytes, ms, bytes / (ms * 1e-3) / 1e9);
    (void)hipFree(buf);
    (void)hipFree(out);
    return rc;
}
