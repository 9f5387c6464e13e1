// Host-side cooking of the triangle meshes that PXG_TRIANGLE_MESH shapes index, as
// TriangleMesh::from_triangles (geometry/triangle_mesh.h:75-170) and MeshBVH::build
// (geometry/mesh_bvh.h:31-111) prepare them:
//
//  * weld: positions closer than eps = 1e-4 x the bounds' diagonal (1e-5 for a single point) share
//    one vertex. Positions are bucketed in cubes of side eps; a position takes the first existing
//    vertex within eps met over the 27 cubes around its own (x offset outermost, z innermost;
//    within a cube in the order of the hash multimap the reference uses), else it starts a vertex.
//    Which vertex a position joins is visible in the triangle indices, so the search order and the
//    container are the reference's;
//  * triangles whose corners weld together, or whose cross product is no longer than 1e-8, are
//    dropped; the rest keep their order;
//  * face normals normalize(cross(b - a, c - a)); neighbours across edges v0-v1, v1-v2, v2-v0 for
//    edges owned by exactly two triangles (the count of edges owned by more is reported);
//  * the triangle tree: a node covers a range of the triangle order, bounded by its triangles'
//    bounds; four or fewer triangles make a leaf, longer ranges split at the median centroid
//    along the bounds' longest axis (std::nth_element with the reference's comparator, so the
//    leaves, and with them the candidate triangles of every query, are the reference's).
//
// Every FP32 expression keeps glm 0.9.9's evaluation order (dot (x*x' + y*y') + z*z', normalize
// v * (1 / sqrt(dot)), min (b < a) ? b : a), and this file is compiled without contraction.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <unordered_map>
#include <utility>
#include <vector>

namespace pxh {

constexpr uint32_t kNoNeighbor = 0xFFFFFFFFu;
constexpr uint32_t kTreeDepthMax = 62;  // levels the GPU walk's stack (64 entries, pxg_mesh.h) covers

struct P3 {
    float x, y, z;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline P3 operator-(P3 a, P3 b) { return P3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 operator+(P3 a, P3 b) { return P3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 operator*(P3 a, float s) { return P3{a.x * s, a.y * s, a.z * s}; }
inline float dot3(P3 a, P3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline P3 cross3(P3 a, P3 b) { return P3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline float len3(P3 a) { return std::sqrt(dot3(a, a)); }
inline float lo1(float a, float b) { return (b < a) ? b : a; }
inline float hi1(float a, float b) { return (a < b) ? b : a; }
inline P3 lo3(P3 a, P3 b) { return P3{lo1(a.x, b.x), lo1(a.y, b.y), lo1(a.z, b.z)}; }
inline P3 hi3(P3 a, P3 b) { return P3{hi1(a.x, b.x), hi1(a.y, b.y), hi1(a.z, b.z)}; }

struct Box3 {
    P3 lo{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
    P3 hi{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest()};
    void grow(const Box3& o) { lo = lo3(lo, o.lo); hi = hi3(hi, o.hi); }
};

struct MeshNodeH {
    Box3 box;
    int32_t left = -1, right = -1;  // children; -1 for a leaf
    uint32_t first = 0, count = 0;  // a leaf's range of the triangle order
};

struct CookedMesh {
    std::vector<P3> vertices;
    std::vector<uint32_t> tris;       // 3 per triangle
    std::vector<P3> normals;          // per triangle
    std::vector<uint32_t> neighbours; // 3 per triangle
    Box3 bounds;
    std::vector<MeshNodeH> nodes;     // depth-first: a node, its left subtree, its right subtree
    std::vector<uint32_t> order;
    uint32_t non_manifold = 0;
    uint32_t depth = 0;               // levels of the tree
    uint32_t n_tris() const { return uint32_t(tris.size() / 3); }
};

inline uint64_t weld_cell(int32_t x, int32_t y, int32_t z) {  // triangle_mesh.h:31-34
    const uint64_t m = 0x1FFFFFu;
    return ((uint64_t(uint32_t(x)) & m) << 42) | ((uint64_t(uint32_t(y)) & m) << 21) | (uint64_t(uint32_t(z)) & m);
}

class TreeBuilder {
public:
    TreeBuilder(CookedMesh& m) : m_(m) {}
    void run() {
        const uint32_t nt = m_.n_tris();
        m_.nodes.clear();
        m_.order.resize(nt);
        m_.depth = 0;
        if (!nt) return;
        box_.resize(nt);
        mid_.resize(nt);
        for (uint32_t t = 0; t < nt; ++t) {
            m_.order[t] = t;
            const P3 a = m_.vertices[m_.tris[3 * t]], b = m_.vertices[m_.tris[3 * t + 1]], c = m_.vertices[m_.tris[3 * t + 2]];
            box_[t].lo = lo3(a, lo3(b, c));
            box_[t].hi = hi3(a, hi3(b, c));
            mid_[t] = (a + b + c) * (1.0f / 3.0f);
        }
        node(0, nt, 1);
    }

private:
    int32_t node(uint32_t first, uint32_t end, uint32_t level) {
        Box3 box;
        for (uint32_t i = first; i < end; ++i) box.grow(box_[m_.order[i]]);
        const int32_t id = int32_t(m_.nodes.size());
        m_.nodes.emplace_back();
        m_.nodes.back().box = box;
        m_.depth = std::max(m_.depth, level);
        if (end - first <= 4) {
            m_.nodes[size_t(id)].first = first;
            m_.nodes[size_t(id)].count = end - first;
            return id;
        }
        const P3 ext = box.hi - box.lo;
        int axis = 0;
        if (ext.y > ext[axis]) axis = 1;
        if (ext.z > ext[axis]) axis = 2;
        const uint32_t half = first + (end - first) / 2;
        const std::vector<P3>& mid = mid_;
        std::nth_element(m_.order.begin() + first, m_.order.begin() + half, m_.order.begin() + end,
                         [&mid, axis](uint32_t a, uint32_t b) { return mid[a][axis] < mid[b][axis]; });
        const int32_t l = node(first, half, level + 1);
        const int32_t r = node(half, end, level + 1);
        m_.nodes[size_t(id)].left = l;
        m_.nodes[size_t(id)].right = r;
        return id;
    }
    CookedMesh& m_;
    std::vector<Box3> box_;
    std::vector<P3> mid_;
};

// False when an index is out of range.
inline bool cook_mesh(const float* pos, uint32_t npos, const uint32_t* ind, uint32_t nind, CookedMesh& out) {
    out = CookedMesh{};
    for (uint32_t i = 0; i < nind; ++i)
        if (ind[i] >= npos) return false;
    if (npos == 0 || nind < 3) return true;
    auto at = [&](uint32_t i) { return P3{pos[3 * size_t(i)], pos[3 * size_t(i) + 1], pos[3 * size_t(i) + 2]}; };

    P3 lo = at(0), hi = at(0);
    for (uint32_t i = 0; i < npos; ++i) {
        lo = lo3(lo, at(i));
        hi = hi3(hi, at(i));
    }
    const float diagonal = len3(hi - lo);
    const float eps = diagonal > 0.0f ? diagonal * 1e-4f : 1e-5f;
    const float side = eps > 0.0f ? eps : 1e-5f;

    std::vector<uint32_t> weld_of(npos);
    std::unordered_multimap<uint64_t, uint32_t> cells;
    cells.reserve(npos);
    for (uint32_t i = 0; i < npos; ++i) {
        const P3 p = at(i);
        const int32_t cx = int32_t(std::floor(p.x / side)), cy = int32_t(std::floor(p.y / side)),
                      cz = int32_t(std::floor(p.z / side));
        uint32_t joined = kNoNeighbor;
        for (int k = 0; k < 27 && joined == kNoNeighbor; ++k) {
            auto range = cells.equal_range(weld_cell(cx + k / 9 - 1, cy + (k / 3) % 3 - 1, cz + k % 3 - 1));
            for (auto it = range.first; it != range.second; ++it)
                if (len3(out.vertices[it->second] - p) <= eps) {
                    joined = it->second;
                    break;
                }
        }
        if (joined == kNoNeighbor) {
            joined = uint32_t(out.vertices.size());
            out.vertices.push_back(p);
            cells.emplace(weld_cell(cx, cy, cz), joined);
        }
        weld_of[i] = joined;
    }

    for (uint32_t t = 0; t + 2 < nind; t += 3) {
        const uint32_t a = weld_of[ind[t]], b = weld_of[ind[t + 1]], c = weld_of[ind[t + 2]];
        if (a == b || b == c || a == c) continue;
        const P3 va = out.vertices[a];
        if (len3(cross3(out.vertices[b] - va, out.vertices[c] - va)) <= 1e-8f) continue;
        out.tris.insert(out.tris.end(), {a, b, c});
    }
    const uint32_t nt = out.n_tris();
    out.normals.resize(nt);
    for (uint32_t t = 0; t < nt; ++t) {
        const P3 a = out.vertices[out.tris[3 * t]];
        const P3 n = cross3(out.vertices[out.tris[3 * t + 1]] - a, out.vertices[out.tris[3 * t + 2]] - a);
        out.normals[t] = n * (1.0f / std::sqrt(dot3(n, n)));
    }

    // edges sorted by their vertex pair: a pair owned twice links its two triangles
    struct Edge { uint64_t key; uint32_t slot; };  // slot = 3 * triangle + edge
    std::vector<Edge> edges(3 * size_t(nt));
    for (uint32_t s = 0; s < 3 * nt; ++s) {
        const uint32_t t = s / 3, e = s % 3;
        uint32_t a = out.tris[3 * t + e], b = out.tris[3 * t + (e + 1) % 3];
        if (b < a) std::swap(a, b);
        edges[s] = Edge{(uint64_t(a) << 32) | b, s};
    }
    std::sort(edges.begin(), edges.end(), [](const Edge& x, const Edge& y) { return x.key < y.key || (x.key == y.key && x.slot < y.slot); });
    out.neighbours.assign(3 * size_t(nt), kNoNeighbor);
    for (size_t i = 0; i < edges.size();) {
        size_t j = i;
        while (j < edges.size() && edges[j].key == edges[i].key) ++j;
        if (j - i == 2) {
            out.neighbours[edges[i].slot] = edges[i + 1].slot / 3;
            out.neighbours[edges[i + 1].slot] = edges[i].slot / 3;
        } else if (j - i > 2) {
            ++out.non_manifold;
        }
        i = j;
    }

    for (const P3& p : out.vertices) {
        out.bounds.lo = lo3(out.bounds.lo, p);
        out.bounds.hi = hi3(out.bounds.hi, p);
    }
    TreeBuilder(out).run();
    return true;
}

}  // namespace pxh
