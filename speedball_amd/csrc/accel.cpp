// accel.cpp — the traversal-only BLAS layout derived from the reference buffers (the SAH
// TLAS of MT_ORDERED mode is built on the device, k_tlas.hip).
//
// It changes no result except which of two EXACTLY equal-t hits wins: instance ids
// stay TLAS slots and triangle ids stay pool indices.
//   * build_node_copies: the threaded BLAS pool in the walk's node format
//     {min.xyz, miss}, {max.xyz, payload}, as one copy in the reference's left-first order
//     or as eight copies where, in copy o, every interior node's children are emitted
//     near-first for ray directions whose sign bits are o (near = smaller projection of
//     the child box centre on the octant direction). Each BLAS keeps its [root, end)
//     range in every copy, so an instance record works unchanged; the walk stays
//     stackless and picks the copy from the local ray direction. perm[i] is the pool node
//     whose bounds copy node i carries (refits re-gather the bounds through it).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "sbgi_host.h"

namespace sbgi {

bool build_node_copies(const uint32_t *nodes, uint32_t n, const std::vector<std::pair<uint32_t, uint32_t>> &roots,
                       uint32_t copies, std::vector<uint32_t> &out, std::vector<uint32_t> &perm) {
    if (copies != 1 && copies != 8) return false;
    out.assign((size_t)copies * n * 8, 0u);
    perm.assign((size_t)copies * n, 0u);
    if (copies == 1) {  // the pool itself, node for node
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t *src = nodes + (size_t)i * 8;
            uint32_t *dst = out.data() + (size_t)i * 8;
            dst[0] = src[0];
            dst[1] = src[1];
            dst[2] = src[2];
            dst[3] = src[6];
            dst[4] = src[3];
            dst[5] = src[4];
            dst[6] = src[5];
            dst[7] = src[7];
            perm[i] = i;
        }
        return true;
    }
    const float *nf = reinterpret_cast<const float *>(nodes);
    for (uint32_t o = 0; o < copies; o++) {
        const float d[3] = {(o & 1) ? -1.f : 1.f, (o & 2) ? -1.f : 1.f, (o & 4) ? -1.f : 1.f};
        uint32_t *oc = out.data() + (size_t)o * n * 8;
        uint32_t *pm = perm.data() + (size_t)o * n;
        for (const auto &re : roots) {
            // iterative pre-order emission; a node's miss is set when its subtree is done
            uint32_t pos = re.first;
            struct Item {
                uint32_t src;
                int32_t finish_of;  // >= 0: finalize marker for output node
            };
            std::vector<Item> work;
            work.push_back({re.first, -1});
            while (!work.empty()) {
                Item it = work.back();
                work.pop_back();
                if (it.finish_of >= 0) {
                    oc[(size_t)it.finish_of * 8 + 3] = pos;
                    continue;
                }
                const uint32_t i = it.src;
                if (i >= n || pos >= re.second) return false;
                const uint32_t me = pos++;
                const uint32_t *src = nodes + (size_t)i * 8;
                uint32_t *dst = oc + (size_t)me * 8;
                dst[0] = src[0];
                dst[1] = src[1];
                dst[2] = src[2];
                dst[4] = src[3];
                dst[5] = src[4];
                dst[6] = src[5];
                dst[7] = src[7];
                pm[me] = i;
                if (src[7] != 0xFFFFFFFFu) {
                    dst[3] = pos;
                    continue;
                }
                const uint32_t l = i + 1;
                if (l >= n) return false;
                const uint32_t r = nodes[(size_t)l * 8 + 6];
                if (r >= n) return false;
                float cl = 0.f, cr = 0.f;
                for (int a = 0; a < 3; a++) {
                    cl += (nf[(size_t)l * 8 + a] + nf[(size_t)l * 8 + 3 + a]) * d[a];
                    cr += (nf[(size_t)r * 8 + a] + nf[(size_t)r * 8 + 3 + a]) * d[a];
                }
                const uint32_t nearc = cl <= cr ? l : r, farc = cl <= cr ? r : l;
                work.push_back({0, (int32_t)me});
                work.push_back({farc, -1});
                work.push_back({nearc, -1});
            }
            if (pos != re.second) return false;
        }
    }
    return true;
}

}  // namespace sbgi
