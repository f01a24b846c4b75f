// scene_builder.cpp — host-side scene assembly in the reference byte layouts.
//
// Replaces buildSpectralScene's acceleration-structure work (js/spectral_scene.js:
// 898-1003 buildLocalBlas/buildTlasRecords, 1188-1240 pool assembly, 1256-1366
// computeDynamic/writeDynamic, 1402-1514 updateTransforms) with native code:
//   * one local-space BLAS per mesh from a binned-SAH builder (leaf <= 8 triangles),
//     flattened to the threaded pre-order layout: left child at i+1, miss = first node
//     after the subtree, leaf payload (count << 24) | pool-absolute triangle offset;
//   * the pool: vertexData stride 8, pool-absolute triIndex, per-triangle material;
//   * the instance table (inverse world 3x4 rows, BLAS range, det sign, active) and a
//     TLAS over instance world AABBs built exactly as buildTlasRecords does (median
//     split by index midpoint on the longest axis, stable centroid sort, leaf <= 2);
//   * per-instance transform updates with the sparse dirty-ancestor TLAS refit of
//     refitFlatTlasDirty (js/gi_refit.js:298-430).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "../../include/sbgi.h"
#include "sbgi_host.h"

namespace sbgi {

namespace {

struct Aabb {
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float *p) {
        for (int a = 0; a < 3; a++) {
            mn[a] = std::min(mn[a], p[a]);
            mx[a] = std::max(mx[a], p[a]);
        }
    }
    void grow(const Aabb &b) {
        for (int a = 0; a < 3; a++) {
            mn[a] = std::min(mn[a], b.mn[a]);
            mx[a] = std::max(mx[a], b.mx[a]);
        }
    }
    float area() const {
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        if (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) return 0.f;
        return 2.f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BuildNode {
    Aabb box;
    uint32_t lo = 0, hi = 0;  // triangle range in the permuted order
    int32_t left = -1, right = -1;
};

constexpr int kBins = 16;
constexpr uint32_t kMaxLeaf = 8;  // three-mesh-bvh targetLeafSize (js/spectral_scene.js:951)
constexpr float kTraversalCost = 1.0f;
constexpr float kIntersectCost = 1.0f;

// Binned SAH over centroids; returns the node index in `nodes`.
int32_t build_rec(std::vector<BuildNode> &nodes, std::vector<uint32_t> &perm, const std::vector<Aabb> &tb,
                  const std::vector<float> &cen, uint32_t lo, uint32_t hi, int depth) {
    BuildNode n;
    n.lo = lo;
    n.hi = hi;
    Aabb cb;
    for (uint32_t i = lo; i < hi; i++) {
        n.box.grow(tb[perm[i]]);
        cb.grow(&cen[3 * perm[i]]);
    }
    int32_t id = (int32_t)nodes.size();
    nodes.push_back(n);
    uint32_t count = hi - lo;
    if (count <= 2 || depth > 60) return id;

    float best_cost = INFINITY;
    int best_axis = -1, best_split = -1;
    for (int axis = 0; axis < 3; axis++) {
        float ext = cb.mx[axis] - cb.mn[axis];
        if (!(ext > 0)) continue;
        Aabb bin_box[kBins];
        uint32_t bin_cnt[kBins] = {0};
        float scale = (float)kBins / ext;
        for (uint32_t i = lo; i < hi; i++) {
            int b = (int)((cen[3 * perm[i] + axis] - cb.mn[axis]) * scale);
            b = std::min(std::max(b, 0), kBins - 1);
            bin_cnt[b]++;
            bin_box[b].grow(tb[perm[i]]);
        }
        float right_area[kBins];
        uint32_t right_cnt[kBins];
        Aabb acc;
        uint32_t c = 0;
        for (int b = kBins - 1; b > 0; b--) {
            acc.grow(bin_box[b]);
            c += bin_cnt[b];
            right_area[b] = acc.area();
            right_cnt[b] = c;
        }
        Aabb lacc;
        uint32_t lc = 0;
        for (int b = 0; b < kBins - 1; b++) {
            lacc.grow(bin_box[b]);
            lc += bin_cnt[b];
            if (lc == 0 || right_cnt[b + 1] == 0) continue;
            float cost = lacc.area() * lc + right_area[b + 1] * right_cnt[b + 1];
            if (cost < best_cost) {
                best_cost = cost;
                best_axis = axis;
                best_split = b;
            }
        }
    }
    float parent_area = n.box.area();
    float leaf_cost = kIntersectCost * count;
    float split_cost = parent_area > 0 ? kTraversalCost + kIntersectCost * best_cost / parent_area : INFINITY;
    uint32_t mid;
    if (best_axis < 0) {
        if (count <= kMaxLeaf) return id;
        mid = lo + count / 2;  // all centroids coincide: split by index
    } else {
        if (count <= kMaxLeaf && leaf_cost <= split_cost) return id;
        float ext = cb.mx[best_axis] - cb.mn[best_axis];
        float scale = (float)kBins / ext;
        auto it = std::stable_partition(perm.begin() + lo, perm.begin() + hi, [&](uint32_t t) {
            int b = (int)((cen[3 * t + best_axis] - cb.mn[best_axis]) * scale);
            b = std::min(std::max(b, 0), kBins - 1);
            return b <= best_split;
        });
        mid = (uint32_t)(it - perm.begin());
        if (mid == lo || mid == hi) mid = lo + count / 2;
    }
    int32_t l = build_rec(nodes, perm, tb, cen, lo, mid, depth + 1);
    int32_t r = build_rec(nodes, perm, tb, cen, mid, hi, depth + 1);
    nodes[id].left = l;
    nodes[id].right = r;
    return id;
}

// Pre-order flatten with miss links (the layout produced by flattenBVHRoot,
// js/spectral_scene.js:658-721).
void flatten(const std::vector<BuildNode> &bn, int32_t id, std::vector<FlatNode> &out) {
    const BuildNode &n = bn[id];
    size_t me = out.size();
    FlatNode f;
    memcpy(f.mn, n.box.mn, sizeof(f.mn));
    memcpy(f.mx, n.box.mx, sizeof(f.mx));
    f.leaf = n.left < 0;
    f.tri_off = n.lo;
    f.tri_cnt = n.hi - n.lo;
    out.push_back(f);
    if (!f.leaf) {
        flatten(bn, n.left, out);
        flatten(bn, n.right, out);
    }
    out[me].miss = (uint32_t)out.size();
}

}  // namespace

// three.js Matrix4.invert (cofactor expansion); zero matrix when det == 0.
void invert4(const double *e, double *o) {
    const double n11 = e[0], n21 = e[1], n31 = e[2], n41 = e[3];
    const double n12 = e[4], n22 = e[5], n32 = e[6], n42 = e[7];
    const double n13 = e[8], n23 = e[9], n33 = e[10], n43 = e[11];
    const double n14 = e[12], n24 = e[13], n34 = e[14], n44 = e[15];
    const double t11 = n23 * n34 * n42 - n24 * n33 * n42 + n24 * n32 * n43 - n22 * n34 * n43 - n23 * n32 * n44 +
                       n22 * n33 * n44;
    const double t12 = n14 * n33 * n42 - n13 * n34 * n42 - n14 * n32 * n43 + n12 * n34 * n43 + n13 * n32 * n44 -
                       n12 * n33 * n44;
    const double t13 = n13 * n24 * n42 - n14 * n23 * n42 + n14 * n22 * n43 - n12 * n24 * n43 - n13 * n22 * n44 +
                       n12 * n23 * n44;
    const double t14 = n14 * n23 * n32 - n13 * n24 * n32 - n14 * n22 * n33 + n12 * n24 * n33 + n13 * n22 * n34 -
                       n12 * n23 * n34;
    const double det = n11 * t11 + n21 * t12 + n31 * t13 + n41 * t14;
    if (det == 0) {
        for (int i = 0; i < 16; i++) o[i] = 0;
        return;
    }
    const double di = 1 / det;
    o[0] = t11 * di;
    o[1] = (n24 * n33 * n41 - n23 * n34 * n41 - n24 * n31 * n43 + n21 * n34 * n43 + n23 * n31 * n44 -
            n21 * n33 * n44) * di;
    o[2] = (n22 * n34 * n41 - n24 * n32 * n41 + n24 * n31 * n42 - n21 * n34 * n42 - n22 * n31 * n44 +
            n21 * n32 * n44) * di;
    o[3] = (n23 * n32 * n41 - n22 * n33 * n41 - n23 * n31 * n42 + n21 * n33 * n42 + n22 * n31 * n43 -
            n21 * n32 * n43) * di;
    o[4] = t12 * di;
    o[5] = (n13 * n34 * n41 - n14 * n33 * n41 + n14 * n31 * n43 - n11 * n34 * n43 - n13 * n31 * n44 +
            n11 * n33 * n44) * di;
    o[6] = (n14 * n32 * n41 - n12 * n34 * n41 - n14 * n31 * n42 + n11 * n34 * n42 + n12 * n31 * n44 -
            n11 * n32 * n44) * di;
    o[7] = (n12 * n33 * n41 - n13 * n32 * n41 + n13 * n31 * n42 - n11 * n33 * n42 - n12 * n31 * n43 +
            n11 * n32 * n43) * di;
    o[8] = t13 * di;
    o[9] = (n14 * n23 * n41 - n13 * n24 * n41 - n14 * n21 * n43 + n11 * n24 * n43 + n13 * n21 * n44 -
            n11 * n23 * n44) * di;
    o[10] = (n12 * n24 * n41 - n14 * n22 * n41 + n14 * n21 * n42 - n11 * n24 * n42 - n12 * n21 * n44 +
             n11 * n22 * n44) * di;
    o[11] = (n13 * n22 * n41 - n12 * n23 * n41 - n13 * n21 * n42 + n11 * n23 * n42 + n12 * n21 * n43 -
             n11 * n22 * n43) * di;
    o[12] = t14 * di;
    o[13] = (n13 * n24 * n31 - n14 * n23 * n31 + n14 * n21 * n33 - n11 * n24 * n33 - n13 * n21 * n34 +
             n11 * n23 * n34) * di;
    o[14] = (n14 * n22 * n31 - n12 * n24 * n31 - n14 * n21 * n32 + n11 * n24 * n32 + n12 * n21 * n34 -
             n11 * n22 * n34) * di;
    o[15] = (n12 * n23 * n31 - n13 * n22 * n31 + n13 * n21 * n32 - n11 * n23 * n32 - n12 * n21 * n33 +
             n11 * n22 * n33) * di;
}

// three.js Matrix4.determinant
double det4(const double *e) {
    const double n11 = e[0], n12 = e[4], n13 = e[8], n14 = e[12];
    const double n21 = e[1], n22 = e[5], n23 = e[9], n24 = e[13];
    const double n31 = e[2], n32 = e[6], n33 = e[10], n34 = e[14];
    const double n41 = e[3], n42 = e[7], n43 = e[11], n44 = e[15];
    return n41 * (+n14 * n23 * n32 - n13 * n24 * n32 - n14 * n22 * n33 + n12 * n24 * n33 + n13 * n22 * n34 -
                  n12 * n23 * n34) +
           n42 * (+n11 * n23 * n34 - n11 * n24 * n33 + n14 * n21 * n33 - n13 * n21 * n34 + n13 * n24 * n31 -
                  n14 * n23 * n31) +
           n43 * (+n11 * n24 * n32 - n11 * n22 * n34 - n14 * n21 * n32 + n12 * n21 * n34 + n14 * n22 * n31 -
                  n12 * n24 * n31) +
           n44 * (-n13 * n22 * n31 - n11 * n23 * n32 + n11 * n22 * n33 + n13 * n21 * n32 - n12 * n21 * n33 +
                  n12 * n23 * n31);
}

// Box3.applyMatrix4: transform the 8 corners (with the projective divide of
// Vector3.applyMatrix4) and take their bounds, in double.
void box_apply(const float *mn, const float *mx, const double *e, double *omn, double *omx) {
    for (int a = 0; a < 3; a++) {
        omn[a] = INFINITY;
        omx[a] = -INFINITY;
    }
    for (int c = 0; c < 8; c++) {
        double x = (c & 1) ? mx[0] : mn[0];
        double y = (c & 2) ? mx[1] : mn[1];
        double z = (c & 4) ? mx[2] : mn[2];
        double w = 1 / (e[3] * x + e[7] * y + e[11] * z + e[15]);
        double p[3] = {(e[0] * x + e[4] * y + e[8] * z + e[12]) * w, (e[1] * x + e[5] * y + e[9] * z + e[13]) * w,
                       (e[2] * x + e[6] * y + e[10] * z + e[14]) * w};
        for (int a = 0; a < 3; a++) {
            omn[a] = std::min(omn[a], p[a]);
            omx[a] = std::max(omx[a], p[a]);
        }
    }
}


// ---- BLAS -----------------------------------------------------------------------
bool build_blas(const float *pos, uint32_t vcount, const uint32_t *idx, uint32_t tcount, std::vector<FlatNode> &out,
                std::vector<uint32_t> &perm, std::string &err) {
    out.clear();
    perm.resize(tcount);
    if (tcount == 0) {
        err = "mesh has no triangles";
        return false;
    }
    std::vector<Aabb> tb(tcount);
    std::vector<float> cen(3 * (size_t)tcount);
    for (uint32_t t = 0; t < tcount; t++) {
        for (int k = 0; k < 3; k++) {
            uint32_t v = idx[3 * (size_t)t + k];
            if (v >= vcount) {
                err = "triangle index out of range";
                return false;
            }
            tb[t].grow(pos + 3 * (size_t)v);
        }
        for (int a = 0; a < 3; a++) cen[3 * (size_t)t + a] = 0.5f * (tb[t].mn[a] + tb[t].mx[a]);
        perm[t] = t;
    }
    std::vector<BuildNode> bn;
    bn.reserve(2 * (size_t)tcount / 2 + 8);
    int32_t root = build_rec(bn, perm, tb, cen, 0, tcount, 0);
    out.reserve(bn.size());
    flatten(bn, root, out);
    return true;
}

// ---- TLAS: buildTlasRecords js/spectral_scene.js:973-1003 ------------------------
void build_tlas(const std::vector<InstAabb> &aabbs, std::vector<TlasRec> &recs, std::vector<uint32_t> &order) {
    const size_t n = aabbs.size();
    order.resize(n);
    for (size_t i = 0; i < n; i++) order[i] = (uint32_t)i;
    recs.clear();
    struct Frame {
        uint32_t lo, hi;
        size_t rec;
        int stage;
    };
    // emit(lo, hi) recursion made explicit: stage 0 = create node and maybe split,
    // stage 1 = children done, set miss.
    std::vector<Frame> st;
    st.push_back({0, (uint32_t)n, 0, 0});
    while (!st.empty()) {
        Frame &f = st.back();
        if (f.stage == 1) {
            recs[f.rec].miss = (uint32_t)recs.size();
            st.pop_back();
            continue;
        }
        TlasRec r;
        for (int a = 0; a < 3; a++) {
            r.bx[a] = INFINITY;
            r.bx[3 + a] = -INFINITY;
        }
        for (uint32_t i = f.lo; i < f.hi; i++) {
            const InstAabb &b = aabbs[order[i]];
            for (int a = 0; a < 3; a++) {
                if (b.mn[a] < r.bx[a]) r.bx[a] = b.mn[a];
                if (b.mx[a] > r.bx[3 + a]) r.bx[3 + a] = b.mx[a];
            }
        }
        f.rec = recs.size();
        f.stage = 1;
        uint32_t lo = f.lo, hi = f.hi;
        if (hi - lo <= 2) {
            r.leaf = true;
            r.off = lo;
            r.cnt = hi - lo;
            recs.push_back(r);
            continue;
        }
        recs.push_back(r);
        double ex = r.bx[3] - r.bx[0], ey = r.bx[4] - r.bx[1], ez = r.bx[5] - r.bx[2];
        int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
        std::stable_sort(order.begin() + lo, order.begin() + hi,
                         [&](uint32_t a, uint32_t b) { return aabbs[a].c[axis] < aabbs[b].c[axis]; });
        uint32_t mid = (lo + hi) >> 1;
        // push right first so the left subtree is emitted next (pre-order)
        st.push_back({mid, hi, 0, 0});
        st.push_back({lo, mid, 0, 0});
    }
}

}  // namespace sbgi

using namespace sbgi;

struct sbgi_scene_builder {
    std::vector<std::vector<float>> materials;
    struct Mesh {
        std::vector<float> vdata;       // stride 8 local
        std::vector<uint32_t> tri;      // permuted, mesh-local vertex ids
        std::vector<uint32_t> tri_mat;  // permuted
        std::vector<FlatNode> nodes;
        float bmn[3], bmx[3];
        uint32_t vert_base = 0, tri_base = 0, node_base = 0;
    };
    std::vector<Mesh> meshes;
    struct Inst {
        uint32_t mesh;
        double world[16];
        bool active;
    };
    std::vector<Inst> insts;
    std::vector<float> lights;  // records in the order they were added
    // finalized arrays
    bool final = false;
    std::vector<float> light_arena;  // analytic records, then emitters
    uint32_t emitter_base = 0, emitter_count = 0;
    std::vector<uint32_t> nodes, tri_index, tri_material, order, slot_of;
    std::vector<float> vertex_data, mats, inst_bounds;
    uint32_t inst_base = 0, tlas_base = 0, tlas_count = 0;
    std::vector<int32_t> parent_by_node, leaf_by_slot;
    std::string err;
};

static void write_instance(sbgi_scene_builder *b, uint32_t slot) {
    const auto &ins = b->insts[b->order[slot]];
    const auto &m = b->meshes[ins.mesh];
    double inv[16];
    invert4(ins.world, inv);
    float *rec = b->mats.data() + b->inst_base + (size_t)slot * 28;
    // column-major -> rows, w = translation (js/spectral_scene.js:1294-1297)
    const int src[12] = {0, 4, 8, 12, 1, 5, 9, 13, 2, 6, 10, 14};
    for (int i = 0; i < 12; i++) rec[i] = (float)inv[src[i]];
    rec[12] = (float)(m.node_base);
    rec[13] = (float)(m.node_base + m.nodes.size());
    rec[14] = det4(ins.world) < 0 ? -1.f : 1.f;
    rec[15] = ins.active ? 1.f : 0.f;
    for (int i = 16; i < 28; i++) rec[i] = 0;
    double mn[3], mx[3];
    if (ins.active) {
        box_apply(m.bmn, m.bmx, ins.world, mn, mx);
    } else {
        for (int a = 0; a < 3; a++) mn[a] = mx[a] = ins.world[12 + a];
    }
    for (int a = 0; a < 3; a++) {
        b->inst_bounds[(size_t)slot * 6 + a] = (float)mn[a];
        b->inst_bounds[(size_t)slot * 6 + 3 + a] = (float)mx[a];
    }
}

extern "C" {

int32_t sbgi_sb_create(sbgi_scene_builder **out) {
    if (!out) return SBGI_E_INVALID;
    *out = new (std::nothrow) sbgi_scene_builder();
    return *out ? SBGI_OK : SBGI_E_NOMEM;
}

int32_t sbgi_sb_add_material(sbgi_scene_builder *b, const float *rec28, uint32_t *id) {
    if (!b || !rec28) return SBGI_E_INVALID;
    b->materials.emplace_back(rec28, rec28 + 28);
    if (id) *id = (uint32_t)b->materials.size() - 1;
    b->final = false;
    return SBGI_OK;
}

int32_t sbgi_sb_add_materials(sbgi_scene_builder *b, const float *recs28, uint32_t n, uint32_t *first_id) {
    if (!b || (n && !recs28)) return SBGI_E_INVALID;
    if (first_id) *first_id = (uint32_t)b->materials.size();
    for (uint32_t i = 0; i < n; i++) b->materials.emplace_back(recs28 + (size_t)i * 28, recs28 + (size_t)i * 28 + 28);
    b->final = false;
    return SBGI_OK;
}

int32_t sbgi_sb_add_mesh(sbgi_scene_builder *b, const float *positions, uint32_t vcount, const float *normals,
                         const float *uvs, const uint32_t *indices, uint32_t tcount, const uint32_t *tri_material,
                         uint32_t *mesh_id) {
    if (!b || !positions || !indices || vcount == 0 || tcount == 0) return SBGI_E_INVALID;
    sbgi_scene_builder::Mesh m;
    std::vector<uint32_t> perm;
    if (!build_blas(positions, vcount, indices, tcount, m.nodes, perm, b->err)) return SBGI_E_INVALID;
    m.vdata.assign((size_t)vcount * 8, 0.f);
    for (uint32_t v = 0; v < vcount; v++) {
        float *d = m.vdata.data() + (size_t)v * 8;
        for (int a = 0; a < 3; a++) d[a] = positions[3 * (size_t)v + a];
        if (normals)
            for (int a = 0; a < 3; a++) d[3 + a] = normals[3 * (size_t)v + a];
        if (uvs)
            for (int a = 0; a < 2; a++) d[6 + a] = uvs[2 * (size_t)v + a];
    }
    m.tri.resize(3 * (size_t)tcount);
    m.tri_mat.resize(tcount);
    for (uint32_t t = 0; t < tcount; t++) {
        uint32_t s = perm[t];
        for (int k = 0; k < 3; k++) m.tri[3 * (size_t)t + k] = indices[3 * (size_t)s + k];
        m.tri_mat[t] = tri_material ? tri_material[s] : 0u;
    }
    memcpy(m.bmn, m.nodes[0].mn, sizeof(m.bmn));
    memcpy(m.bmx, m.nodes[0].mx, sizeof(m.bmx));
    b->meshes.push_back(std::move(m));
    if (mesh_id) *mesh_id = (uint32_t)b->meshes.size() - 1;
    b->final = false;
    return SBGI_OK;
}

int32_t sbgi_sb_add_instance(sbgi_scene_builder *b, uint32_t mesh_id, const double *world16, int32_t active,
                             uint32_t *instance_id) {
    if (!b || !world16 || mesh_id >= b->meshes.size()) return SBGI_E_INVALID;
    sbgi_scene_builder::Inst in;
    in.mesh = mesh_id;
    memcpy(in.world, world16, sizeof(in.world));
    in.active = active != 0;
    b->insts.push_back(in);
    if (instance_id) *instance_id = (uint32_t)b->insts.size() - 1;
    b->final = false;
    return SBGI_OK;
}

int32_t sbgi_sb_add_light(sbgi_scene_builder *b, const float *rec17) {
    if (!b || !rec17) return SBGI_E_INVALID;
    if (b->lights.size() / 17 >= SBGI_MAX_LIGHTS) {
        b->err = "more than MAX_LIGHTS (64) lights";
        return SBGI_E_CAPACITY;
    }
    b->lights.insert(b->lights.end(), rec17, rec17 + 17);
    b->final = false;
    return SBGI_OK;
}

int32_t sbgi_sb_finalize(sbgi_scene_builder *b, sbgi_scene_view *view) {
    if (!b || !view) return SBGI_E_INVALID;
    if (b->meshes.empty() || b->insts.empty()) {
        b->err = "scene needs at least one mesh and one instance";
        return SBGI_E_INVALID;
    }
    if (b->materials.empty()) {
        // MeshStandardMaterial defaults through materialToUber (js/spectral_scene.js:214-307)
        std::vector<float> d(28, 0.f);
        d[0] = d[1] = d[2] = 1.f;
        d[3] = 1.f;
        d[6] = 1.5f;
        d[10] = 1.f;
        for (int i = 12; i <= 16; i++) d[i] = -1.f;
        d[17] = 1.f;
        d[18] = d[19] = 1.f;
        d[24] = -1.f;
        d[25] = -1.f;
        d[26] = 6.f;
        b->materials.push_back(d);
    }
    uint64_t pv = 0, pt = 0, pn = 0;
    for (auto &m : b->meshes) {
        m.vert_base = (uint32_t)pv;
        m.tri_base = (uint32_t)pt;
        m.node_base = (uint32_t)pn;
        pv += m.vdata.size() / 8;
        pt += m.tri_mat.size();
        pn += m.nodes.size();
    }
    if (pt >= (1u << 24) || pn >= (1u << 24)) {
        b->err = "BVH pool exceeds the 24-bit index ceiling";
        return SBGI_E_CAPACITY;
    }
    if (b->insts.size() >= (1u << 24)) {
        b->err = "too many instances for the TLAS leaf payload";
        return SBGI_E_CAPACITY;
    }
    for (auto &m : b->meshes)
        for (uint32_t tm : m.tri_mat)
            if (tm >= b->materials.size()) {
                b->err = "triangle material index out of range";
                return SBGI_E_INVALID;
            }
    b->vertex_data.assign(pv * 8, 0.f);
    b->tri_index.assign(pt * 3, 0);
    b->tri_material.assign(pt, 0);
    b->nodes.assign(pn * 8, 0);
    for (auto &m : b->meshes) {
        memcpy(b->vertex_data.data() + (size_t)m.vert_base * 8, m.vdata.data(), m.vdata.size() * 4);
        for (size_t t = 0; t < m.tri_mat.size(); t++) {
            for (int k = 0; k < 3; k++) b->tri_index[(m.tri_base + t) * 3 + k] = m.tri[t * 3 + k] + m.vert_base;
            b->tri_material[m.tri_base + t] = m.tri_mat[t];
        }
        for (size_t i = 0; i < m.nodes.size(); i++) {
            const FlatNode &f = m.nodes[i];
            uint32_t *o = b->nodes.data() + (m.node_base + i) * 8;
            memcpy(o, f.mn, 12);
            memcpy(o + 3, f.mx, 12);
            o[6] = f.miss + m.node_base;
            o[7] = f.leaf ? (((f.tri_cnt & 0xFFu) << 24) | ((f.tri_off + m.tri_base) & 0x00FFFFFFu)) : 0xFFFFFFFFu;
        }
    }
    // dynamic tail (computeDynamic + buildTlasRecords + writeDynamic)
    const size_t ni = b->insts.size();
    std::vector<InstAabb> aabbs(ni);
    for (size_t i = 0; i < ni; i++) {
        const auto &ins = b->insts[i];
        const auto &m = b->meshes[ins.mesh];
        if (ins.active) {
            box_apply(m.bmn, m.bmx, ins.world, aabbs[i].mn, aabbs[i].mx);
        } else {
            for (int a = 0; a < 3; a++) aabbs[i].mn[a] = aabbs[i].mx[a] = ins.world[12 + a];
        }
        for (int a = 0; a < 3; a++) aabbs[i].c[a] = (aabbs[i].mn[a] + aabbs[i].mx[a]) * 0.5;
    }
    std::vector<TlasRec> recs;
    build_tlas(aabbs, recs, b->order);
    b->slot_of.assign(ni, 0);
    for (size_t s = 0; s < ni; s++) b->slot_of[b->order[s]] = (uint32_t)s;
    b->inst_base = (uint32_t)(b->materials.size() * 28);
    b->tlas_base = b->inst_base + (uint32_t)ni * 28;
    b->tlas_count = (uint32_t)recs.size();
    b->mats.assign((size_t)b->tlas_base + recs.size() * 12, 0.f);
    for (size_t i = 0; i < b->materials.size(); i++) memcpy(b->mats.data() + i * 28, b->materials[i].data(), 28 * 4);
    b->inst_bounds.assign(ni * 6, 0.f);
    for (uint32_t s = 0; s < ni; s++) write_instance(b, s);
    for (size_t i = 0; i < recs.size(); i++) {
        float *o = b->mats.data() + b->tlas_base + i * 12;
        for (int a = 0; a < 6; a++) o[a] = (float)recs[i].bx[a];
        o[6] = (float)recs[i].miss;
        o[7] = recs[i].leaf ? (float)recs[i].off : 0.f;
        o[8] = recs[i].leaf ? (float)recs[i].cnt : 0.f;
    }
    // createFlatTlasRefitIndex (js/gi_refit.js:298-346)
    b->parent_by_node.assign(recs.size(), -1);
    b->leaf_by_slot.assign(ni, -1);
    for (size_t i = 0; i < recs.size(); i++) {
        if (recs[i].leaf) {
            for (uint32_t s = recs[i].off; s < recs[i].off + recs[i].cnt; s++) b->leaf_by_slot[s] = (int32_t)i;
        } else {
            size_t l = i + 1, r = recs[l].miss;
            b->parent_by_node[l] = (int32_t)i;
            b->parent_by_node[r] = (int32_t)i;
        }
    }
    b->final = true;
    view->bvh_nodes = b->nodes.data();
    view->node_count = (uint32_t)pn;
    view->tri_count = (uint32_t)pt;
    view->tri_index = b->tri_index.data();
    view->vertex_data = b->vertex_data.data();
    view->vertex_count = (uint32_t)pv;
    view->materials_len = (uint32_t)b->mats.size();
    view->tri_material = b->tri_material.data();
    view->materials = b->mats.data();
    view->inst_base = b->inst_base;
    view->inst_count = (uint32_t)ni;
    view->tlas_base = b->tlas_base;
    view->tlas_node_count = b->tlas_count;
    // The light arena of selectGiLightRecords (js/gi_probes.js:4273-4289, the order-keeping
    // selection of the non-clustered arena): analytic lights (type < 2.5) in the order added,
    // up to 64 minus the emitters, then the emitters (type 3), at most 16; other records
    // are dropped.
    std::vector<const float *> analytic, emitters;
    for (size_t i = 0; i + 17 <= b->lights.size(); i += 17) {
        const float *r = b->lights.data() + i;
        if (std::fabs(r[0] - 3.f) < 0.5f) emitters.push_back(r);
        else if (r[0] < 2.5f) analytic.push_back(r);
    }
    b->emitter_count = (uint32_t)std::min<size_t>(emitters.size(), SBGI_MAX_EMITTERS);
    const size_t n_analytic = std::min<size_t>(analytic.size(), SBGI_MAX_LIGHTS - b->emitter_count);
    b->light_arena.clear();
    for (size_t i = 0; i < n_analytic; i++) b->light_arena.insert(b->light_arena.end(), analytic[i], analytic[i] + 17);
    for (size_t i = 0; i < b->emitter_count; i++)
        b->light_arena.insert(b->light_arena.end(), emitters[i], emitters[i] + 17);
    b->emitter_base = (uint32_t)n_analytic;
    view->lights = b->light_arena.empty() ? nullptr : b->light_arena.data();
    view->light_count = (uint32_t)(b->light_arena.size() / 17);
    view->emitter_base = b->emitter_base;
    view->emitter_count = b->emitter_count;
    view->pad = 0;
    return SBGI_OK;
}

int32_t sbgi_sb_tlas_order(const sbgi_scene_builder *b, uint32_t *order, uint32_t n) {
    if (!b || !order || !b->final || n != b->order.size()) return SBGI_E_INVALID;
    memcpy(order, b->order.data(), n * sizeof(uint32_t));
    return SBGI_OK;
}

int32_t sbgi_sb_mesh_ranges(const sbgi_scene_builder *b, uint32_t mesh_id, uint32_t *out6) {
    if (!b || !out6 || !b->final || mesh_id >= b->meshes.size()) return SBGI_E_INVALID;
    const auto &m = b->meshes[mesh_id];
    out6[0] = m.vert_base;
    out6[1] = (uint32_t)(m.vdata.size() / 8);
    out6[2] = m.tri_base;
    out6[3] = (uint32_t)m.tri_mat.size();
    out6[4] = m.node_base;
    out6[5] = m.node_base + (uint32_t)m.nodes.size();
    return SBGI_OK;
}

int32_t sbgi_sb_set_transform(sbgi_scene_builder *b, uint32_t instance_id, const double *world16, uint32_t *slot) {
    if (!b || !world16 || instance_id >= b->insts.size()) return SBGI_E_INVALID;
    if (!b->final) {
        b->err = "scene not finalized";
        return SBGI_E_STATE;
    }
    memcpy(b->insts[instance_id].world, world16, 16 * sizeof(double));
    uint32_t s = b->slot_of[instance_id];
    write_instance(b, s);
    if (slot) *slot = s;
    // refitFlatTlasDirty: recompute the leaf owning the slot and its ancestors.
    float *tl = b->mats.data() + b->tlas_base;
    for (int32_t node = b->leaf_by_slot[s]; node >= 0; node = b->parent_by_node[node]) {
        float *nd = tl + (size_t)node * 12;
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (nd[8] > 0) {
            uint32_t off = (uint32_t)nd[7], cnt = (uint32_t)nd[8];
            for (uint32_t q = off; q < off + cnt; q++) {
                const float *bb = b->inst_bounds.data() + (size_t)q * 6;
                for (int a = 0; a < 3; a++) {
                    if (bb[a] < mn[a]) mn[a] = bb[a];
                    if (bb[3 + a] > mx[a]) mx[a] = bb[3 + a];
                }
            }
        } else {
            size_t l = (size_t)node + 1, r = (size_t)tl[l * 12 + 6];
            for (int a = 0; a < 3; a++) {
                mn[a] = std::min(tl[l * 12 + a], tl[r * 12 + a]);
                mx[a] = std::max(tl[l * 12 + 3 + a], tl[r * 12 + 3 + a]);
            }
        }
        for (int a = 0; a < 3; a++) {
            nd[a] = mn[a];
            nd[3 + a] = mx[a];
        }
    }
    return SBGI_OK;
}

const char *sbgi_sb_last_error(const sbgi_scene_builder *b) { return b ? b->err.c_str() : "null builder"; }

int32_t sbgi_sb_destroy(sbgi_scene_builder *b) {
    delete b;
    return SBGI_OK;
}

}  // extern "C"
