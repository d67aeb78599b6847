// Native static-geometry packer (host C++): the triangle BVH and the sample-cluster
// tree of the packed mesh layout, one template per call.  The splits follow the
// reference packer (src/sand_atlas_dem/mesh.py:28-70 `partition`, 160-202 `build`):
// preorder, the widest FP32 extent (first axis on ties), a stable sort on that
// axis, the lower half first; bounds rounded outward by one FP32 step.  Each call
// works on its own buffers, so templates can be packed from several threads.
#include "sand_dem_pack.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
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

// ---------------------------------------------------------------- SDF grids
// Exact unsigned distance to the nearest triangle (double precision, BVH with box
// pruning), signed by the angle-weighted pseudonormal of the nearest feature
// (Baerentzen & Aanaes 2005): negative inside.
struct D3 { double x, y, z; };
inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 mul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Closest point of triangle abc to p and its feature: face 0, edges AB/BC/CA 1/2/3,
// vertices A/B/C 4/5/6 (Voronoi-region tests, Ericson, Real-Time Collision
// Detection, 5.1.5).
D3 closest(D3 p, D3 a, D3 b, D3 c, int& feature) {
    D3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { feature = 4; return a; }
    D3 bp = sub(p, b);
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { feature = 5; return b; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) { feature = 1; return add(a, mul(ab, d1 / (d1 - d3))); }
    D3 cp = sub(p, c);
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { feature = 6; return c; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) { feature = 3; return add(a, mul(ac, d2 / (d2 - d6))); }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        feature = 2;
        return add(b, mul(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    }
    double denom = 1.0 / (va + vb + vc);
    feature = 0;
    return add(a, add(mul(ab, vb * denom), mul(ac, vc * denom)));
}

struct SdfTree {
    struct Node { double lo[3], hi[3]; uint32_t first, count, right; };
    const double* tri;
    std::vector<uint32_t> ids;
    std::vector<Node> nodes;

    uint32_t build(uint32_t begin, uint32_t end) {
        const uint32_t node = uint32_t(nodes.size());
        nodes.push_back({});
        Node n{};
        for (int a = 0; a < 3; a++) n.lo[a] = INFINITY, n.hi[a] = -INFINITY;
        for (uint32_t i = begin; i < end; i++)
            for (int v = 0; v < 3; v++)
                for (int a = 0; a < 3; a++) {
                    double x = tri[9 * size_t(ids[i]) + 3 * v + a];
                    n.lo[a] = std::min(n.lo[a], x);
                    n.hi[a] = std::max(n.hi[a], x);
                }
        n.first = begin;
        n.count = end - begin;
        if (end - begin > 4) {
            int axis = 0;
            for (int a = 1; a < 3; a++)
                if (n.hi[a] - n.lo[a] > n.hi[axis] - n.lo[axis]) axis = a;
            auto key = [&](uint32_t t) {
                const double* q = tri + 9 * size_t(t);
                return q[axis] + q[3 + axis] + q[6 + axis];
            };
            std::stable_sort(ids.begin() + begin, ids.begin() + end, [&](uint32_t x, uint32_t y) { return key(x) < key(y); });
            const uint32_t mid = begin + (end - begin) / 2;
            n.count = 0;  // internal: left child is node + 1
            nodes[node] = n;
            build(begin, mid);
            n.right = build(mid, end);
        }
        nodes[node] = n;
        return node;
    }

    static double boxDistance2(const Node& n, D3 p) {
        double q[3] = {p.x, p.y, p.z}, sum = 0;
        for (int a = 0; a < 3; a++) {
            double d = std::max(std::max(n.lo[a] - q[a], q[a] - n.hi[a]), 0.0);
            sum += d * d;
        }
        return sum;
    }

    // Nearest triangle (first in visit order at the minimum) and its closest point.
    void query(D3 p, double& best, uint32_t& triangle, D3& point, int& feature) const {
        uint32_t stack[64];
        int top = 0;
        stack[top++] = 0;
        best = INFINITY;
        while (top) {
            const Node& n = nodes[stack[--top]];
            if (boxDistance2(n, p) > best) continue;
            if (n.count) {
                for (uint32_t i = n.first; i < n.first + n.count; i++) {
                    const double* q = tri + 9 * size_t(ids[i]);
                    int f;
                    D3 c = closest(p, {q[0], q[1], q[2]}, {q[3], q[4], q[5]}, {q[6], q[7], q[8]}, f);
                    D3 d = sub(p, c);
                    double d2 = dot(d, d);
                    if (d2 < best) best = d2, triangle = ids[i], point = c, feature = f;
                }
                continue;
            }
            const uint32_t left = uint32_t(&n - nodes.data()) + 1, right = n.right;
            double dl = boxDistance2(nodes[left], p), dr = boxDistance2(nodes[right], p);
            if (dl <= dr) { stack[top++] = right; stack[top++] = left; }
            else { stack[top++] = left; stack[top++] = right; }
        }
    }
};

}  // namespace

extern "C" {

int sad_sdf_grid(const double* triangles, const double* normals, uint32_t count, const double* origin,
                 double spacing, const uint32_t* dims, uint32_t threads, float* out) {
    if (!triangles || !normals || !origin || !dims || !out || count == 0 || !(spacing > 0)) return ARGUMENT;
    if (dims[0] == 0 || dims[1] == 0 || dims[2] == 0) return ARGUMENT;
    SdfTree tree;
    tree.tri = triangles;
    tree.ids.resize(count);
    for (uint32_t i = 0; i < count; i++) tree.ids[i] = i;
    tree.build(0, count);
    const uint32_t nx = dims[0], ny = dims[1], nz = dims[2];
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (uint32_t x; (x = next.fetch_add(1)) < nx;)
            for (uint32_t y = 0; y < ny; y++)
                for (uint32_t z = 0; z < nz; z++) {
                    D3 p = {origin[0] + spacing * double(x), origin[1] + spacing * double(y),
                            origin[2] + spacing * double(z)};
                    double best;
                    uint32_t t = 0;
                    D3 c{};
                    int feature = 0;
                    tree.query(p, best, t, c, feature);
                    const double* n = normals + 21 * size_t(t) + 3 * feature;
                    double side = dot(sub(p, c), {n[0], n[1], n[2]}) < 0 ? -1.0 : 1.0;
                    out[(size_t(x) * ny + y) * nz + z] = float(side * std::sqrt(best));
                }
    };
    const uint32_t workers = std::max(1u, std::min(threads, nx));
    std::vector<std::thread> pool;
    for (uint32_t k = 1; k < workers; k++) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    return OK;
}


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
