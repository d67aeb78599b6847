// Native static-geometry packer (host C++): the triangle BVH and the sample-cluster
// tree of the packed mesh layout, one template per call.  The splits follow the
// reference packer (src/sand_atlas_dem/mesh.py:28-70 `partition`, 160-202 `build`):
// preorder, the widest FP32 extent (first axis on ties), a stable sort on that
// axis, the lower half first; bounds rounded outward by one FP32 step.  Each call
// works on its own buffers, so templates can be packed from several threads.
#include "sand_dem_pack.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

constexpr int OK = 0, ARGUMENT = 2, TOO_DEEP = 5;
constexpr uint32_t MAX_DEPTH = 30;

uint32_t nodes(uint32_t items, uint32_t leaf) {
    if (items <= leaf) return 1;
    return 1 + nodes(items / 2, leaf) + nodes(items - items / 2, leaf);
}

// Index of the widest extent (max - min in FP32); the first one on ties.
int widest(const float* lo, const float* hi) {
    int axis = 0;
    float best = hi[0] - lo[0];
    for (int a = 1; a < 3; a++) {
        float extent = hi[a] - lo[a];
        if (extent > best) best = extent, axis = a;
    }
    return axis;
}

void writeBounds(float* lower, float* upper, uint32_t row, const float* lo, const float* hi) {
    for (int a = 0; a < 3; a++) {
        lower[4 * row + a] = std::nextafter(lo[a], -INFINITY);
        upper[4 * row + a] = std::nextafter(hi[a], INFINITY);
    }
    lower[4 * row + 3] = 0.0f;
    upper[4 * row + 3] = 0.0f;
}

struct Bvh {
    const float* tri;
    const float* centre;
    uint32_t nodeBase, triangleBase;
    uint32_t *links, *order;
    float *lower, *upper;
    uint32_t next = 0, packed = 0, depth = 0;

    void build(uint32_t* ids, uint32_t n, uint32_t level) {
        const uint32_t node = next++;
        depth = std::max(depth, level);
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = 0; i < n; i++)
            for (int v = 0; v < 3; v++)
                for (int a = 0; a < 3; a++) {
                    float x = tri[9 * size_t(ids[i]) + 3 * v + a];
                    lo[a] = std::min(lo[a], x);
                    hi[a] = std::max(hi[a], x);
                }
        writeBounds(lower, upper, node, lo, hi);
        uint32_t* link = links + 4 * size_t(node);
        link[0] = link[1] = link[2] = link[3] = 0;
        if (n <= 4) {
            link[1] = triangleBase + packed;
            link[2] = n;
            for (uint32_t i = 0; i < n; i++) order[packed++] = ids[i];
        } else {
            float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t i = 0; i < n; i++)
                for (int a = 0; a < 3; a++) {
                    float c = centre[3 * size_t(ids[i]) + a];
                    clo[a] = std::min(clo[a], c);
                    chi[a] = std::max(chi[a], c);
                }
            const int axis = widest(clo, chi);
            std::stable_sort(ids, ids + n, [&](uint32_t x, uint32_t y) {
                return centre[3 * size_t(x) + axis] < centre[3 * size_t(y) + axis];
            });
            build(ids, n / 2, level + 1);
            build(ids + n / 2, n - n / 2, level + 1);
        }
        link[0] = nodeBase + next;  // escape: the node after this subtree (leaves too)
    }
};

struct Clusters {
    const float* point;
    uint32_t *ranges, *keys;
    float *lower, *upper;
    uint32_t next = 0, written = 0, keyBase = 0, clusterBase = 0;

    void split(uint32_t* ids, uint32_t n) {
        const uint32_t node = next++;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = 0; i < n; i++)
            for (int a = 0; a < 3; a++) {
                float x = point[3 * size_t(ids[i]) + a];
                lo[a] = std::min(lo[a], x);
                hi[a] = std::max(hi[a], x);
            }
        writeBounds(lower, upper, node, lo, hi);
        uint32_t* range = ranges + 4 * size_t(node);
        range[0] = range[1] = range[2] = range[3] = 0;
        if (n > 64) {
            const int axis = widest(lo, hi);
            std::stable_sort(ids, ids + n, [&](uint32_t x, uint32_t y) {
                return point[3 * size_t(x) + axis] < point[3 * size_t(y) + axis];
            });
            split(ids, n / 2);
            split(ids + n / 2, n - n / 2);
        } else {
            range[0] = keyBase + written;
            range[1] = n;  // a zero count marks an internal node
            for (uint32_t i = 0; i < n; i++) keys[written++] = ids[i];
        }
        range[2] = clusterBase + next;
    }
};

}  // namespace

extern "C" {

uint32_t sad_pack_node_count(uint32_t items, uint32_t leaf_size) {
    if (items == 0 || leaf_size == 0) return 0;
    return nodes(items, leaf_size);
}

int sad_pack_bvh(const float* triangles, const float* centres, uint32_t count, uint32_t node_base,
                 uint32_t triangle_base, uint32_t* links, float* lower, float* upper, uint32_t* order,
                 uint32_t* depth) {
    if (!triangles || !centres || !links || !lower || !upper || !order || !depth || count == 0) return ARGUMENT;
    std::vector<uint32_t> ids(count);
    for (uint32_t i = 0; i < count; i++) ids[i] = i;
    Bvh b{triangles, centres, node_base, triangle_base, links, order, lower, upper};
    b.build(ids.data(), count, 0);
    *depth = b.depth;
    return b.depth > MAX_DEPTH ? TOO_DEEP : OK;
}

int sad_pack_clusters(const float* points, uint32_t count, uint32_t cluster_base, uint32_t key_base,
                      uint32_t* ranges, float* lower, float* upper, uint32_t* keys, uint32_t* frontier) {
    if (!points || !ranges || !lower || !upper || !keys || !frontier || count == 0) return ARGUMENT;
    std::vector<uint32_t> ids(count);
    for (uint32_t i = 0; i < count; i++) ids[i] = i;
    Clusters c{points, ranges, keys, lower, upper};
    c.keyBase = key_base;
    c.clusterBase = cluster_base;
    c.split(ids.data(), count);
    // Up to 32 disjoint complete subtrees: repeatedly replace the largest internal
    // one (the first in list order on ties) by its two children, in place.
    auto escape = [&](uint32_t local) { return ranges[4 * size_t(local) + 2] - cluster_base; };
    std::vector<uint32_t> pending{0};
    while (pending.size() < 32) {
        int pick = -1;
        uint32_t size = 0;
        for (size_t k = 0; k < pending.size(); k++) {
            uint32_t node = pending[k];
            if (ranges[4 * size_t(node) + 1] != 0) continue;
            uint32_t extent = escape(node) - node;
            if (pick < 0 || extent > size) pick = int(k), size = extent;
        }
        if (pick < 0) break;
        const uint32_t left = pending[pick] + 1;
        pending[pick] = left;
        pending.insert(pending.begin() + pick + 1, escape(left));
    }
    for (uint32_t k = 0; k < 32; k++) {
        uint32_t* row = frontier + 4 * k;
        row[0] = row[1] = row[2] = row[3] = 0;
        if (k < pending.size()) row[0] = cluster_base + pending[k], row[1] = cluster_base + escape(pending[k]);
    }
    return OK;
}

}  // extern "C"
