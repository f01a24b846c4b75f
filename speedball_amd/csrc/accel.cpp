// accel.cpp — traversal-only acceleration layouts derived from the reference buffers.
//
// Neither changes any result except which of two EXACTLY equal-t hits wins: instance ids
// stay TLAS slots and triangle ids stay pool indices.
//   * build_sah_tlas: a binned-SAH TLAS over the instance world AABBs with one slot per
//     leaf (so any partition can reference the fixed slot order), in the reference's
//     12-float node format. The reference TLAS is a median split on the longest axis
//     (js/spectral_scene.js:973-1003); with large overlapping instances the SAH tree
//     enters ~2.5x fewer instances per ray.
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

namespace {

struct Box {
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float *b6) {
        for (int a = 0; a < 3; a++) {
            mn[a] = std::min(mn[a], b6[a]);
            mx[a] = std::max(mx[a], b6[3 + a]);
        }
    }
    void grow(const Box &b) {
        for (int a = 0; a < 3; a++) {
            mn[a] = std::min(mn[a], b.mn[a]);
            mx[a] = std::max(mx[a], b.mx[a]);
        }
    }
    double area() const {
        double d[3];
        for (int a = 0; a < 3; a++) {
            d[a] = (double)mx[a] - (double)mn[a];
            if (!(d[a] >= 0)) return 0.0;
        }
        return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
    }
};

void emit_sah(const float *bounds, std::vector<uint32_t> &ids, size_t lo, size_t hi, std::vector<float> &out) {
    const size_t me = out.size() / 12;
    out.resize(out.size() + 12, 0.f);
    Box b, cb;
    for (size_t i = lo; i < hi; i++) {
        const float *bb = bounds + (size_t)ids[i] * 6;
        b.grow(bb);
        float c[6];
        for (int a = 0; a < 3; a++) c[a] = c[3 + a] = 0.5f * (bb[a] + bb[3 + a]);
        cb.grow(c);
    }
    float *rec = out.data() + me * 12;
    for (int a = 0; a < 3; a++) {
        rec[a] = b.mn[a];
        rec[3 + a] = b.mx[a];
    }
    if (hi - lo == 1) {
        rec[7] = (float)ids[lo];
        rec[8] = 1.f;
        rec[6] = (float)(me + 1);
        return;
    }
    constexpr int kBins = 32;
    double best = INFINITY;
    int best_axis = -1, best_bin = -1;
    for (int axis = 0; axis < 3; axis++) {
        const float ext = cb.mx[axis] - cb.mn[axis];
        if (!(ext > 0)) continue;
        Box bins[kBins];
        uint32_t cnt[kBins] = {0};
        for (size_t i = lo; i < hi; i++) {
            const float *bb = bounds + (size_t)ids[i] * 6;
            const float c = 0.5f * (bb[axis] + bb[3 + axis]);
            int k = (int)((c - cb.mn[axis]) / ext * kBins);
            k = std::min(std::max(k, 0), kBins - 1);
            cnt[k]++;
            bins[k].grow(bb);
        }
        double ra[kBins];
        uint32_t rc[kBins];
        Box acc;
        uint32_t c = 0;
        for (int k = kBins - 1; k > 0; k--) {
            acc.grow(bins[k]);
            c += cnt[k];
            ra[k] = acc.area();
            rc[k] = c;
        }
        Box lacc;
        uint32_t lc = 0;
        for (int k = 0; k < kBins - 1; k++) {
            lacc.grow(bins[k]);
            lc += cnt[k];
            if (!lc || !rc[k + 1]) continue;
            const double cost = lacc.area() * lc + ra[k + 1] * rc[k + 1];
            if (cost < best) {
                best = cost;
                best_axis = axis;
                best_bin = k;
            }
        }
    }
    size_t mid;
    if (best_axis < 0) {
        mid = lo + (hi - lo) / 2;
    } else {
        const float ext = cb.mx[best_axis] - cb.mn[best_axis];
        auto it = std::stable_partition(ids.begin() + lo, ids.begin() + hi, [&](uint32_t id) {
            const float *bb = bounds + (size_t)id * 6;
            const float c = 0.5f * (bb[best_axis] + bb[3 + best_axis]);
            int k = (int)((c - cb.mn[best_axis]) / ext * kBins);
            k = std::min(std::max(k, 0), kBins - 1);
            return k <= best_bin;
        });
        mid = (size_t)(it - ids.begin());
        if (mid == lo || mid == hi) mid = lo + (hi - lo) / 2;
    }
    emit_sah(bounds, ids, lo, mid, out);
    emit_sah(bounds, ids, mid, hi, out);
    out[me * 12 + 6] = (float)(out.size() / 12);
}

}  // namespace

void build_sah_tlas(const float *inst_bounds, uint32_t n, std::vector<float> &out) {
    out.clear();
    if (n == 0) return;
    std::vector<uint32_t> ids(n);
    for (uint32_t i = 0; i < n; i++) ids[i] = i;
    out.reserve((size_t)(2 * n) * 12);
    emit_sah(inst_bounds, ids, 0, n, out);
}

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
