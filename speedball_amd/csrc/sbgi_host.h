// sbgi_host.h — host-side types shared by the scene builder and the field controller.
#pragma once
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace sbgi {

// One node of a flattened BLAS before pool offsets are applied.
struct FlatNode {
    float mn[3], mx[3];
    uint32_t miss = 0;
    bool leaf = false;
    uint32_t tri_off = 0, tri_cnt = 0;
};

// Instance world AABB in double, as computeDynamic builds it (js/spectral_scene.js:1308-1312).
struct InstAabb {
    double mn[3], mx[3], c[3];
};

// One TLAS record of buildTlasRecords (js/spectral_scene.js:973-1003).
struct TlasRec {
    double bx[6];
    bool leaf = false;
    uint32_t off = 0, cnt = 0, miss = 0;
};

// three.js Matrix4.invert / determinant / Box3.applyMatrix4 on column-major doubles
void invert4(const double *e, double *o);
double det4(const double *e);
void box_apply(const float *mn, const float *mx, const double *e, double *omn, double *omx);

// traversal-only layouts (accel.cpp)
bool build_node_copies(const uint32_t *nodes, uint32_t n, const std::vector<std::pair<uint32_t, uint32_t>> &roots,
                       uint32_t copies, std::vector<uint32_t> &out, std::vector<uint32_t> &perm);

bool build_blas(const float *pos, uint32_t vcount, const uint32_t *idx, uint32_t tcount,
                std::vector<FlatNode> &out, std::vector<uint32_t> &perm, std::string &err);
void build_tlas(const std::vector<InstAabb> &aabbs, std::vector<TlasRec> &recs, std::vector<uint32_t> &order);

}  // namespace sbgi
