// ORACLE -- test infrastructure only. Nothing in the product may include, link or run this.
//
// Static triangle meshes: the cooking of TriangleMesh::from_triangles (geometry/triangle_mesh.h:
// 75-170: weld through a spatial hash, drop degenerate triangles, face normals, edge adjacency),
// the MeshBVH the contact path narrows triangles with (geometry/mesh_bvh.h:31-111), and the
// convex-against-mesh manifold of collision/mesh_contact.h (sphere and capsule per triangle with
// the internal-edge correction and embedded recovery; the box and hull manifolds of corners over
// triangles and triangle vertices inside the shape, reduced to four well-spread points), restated
// function by function. Cooking is FP32 as in the reference; the contact functions run in T.
#pragma once
#include "pxo_world.hpp"

#include <array>
#include <unordered_map>
#include <vector>

namespace pxo {

constexpr uint32_t k_no_neighbor = 0xFFFFFFFFu;  // triangle_mesh.h:38

struct MeshNode {  // mesh_bvh.h:82-89
    AABB<float> b;
    int32_t left = -1, right = -1;
    uint32_t start = 0, count = 0;
};

// geometry::TriangleMesh built by from_triangles: welded vertices, triangle indices, unit face
// normals, per-edge neighbour triangles (edge 0: v0-v1, 1: v1-v2, 2: v2-v0), bounds of every
// welded vertex, and the BVH over the triangles.
struct Mesh {
    std::vector<V3<float>> v;
    std::vector<uint32_t> idx;
    std::vector<V3<float>> n;
    std::vector<std::array<uint32_t, 3>> adj;
    AABB<float> bounds;
    std::vector<MeshNode> nodes;
    std::vector<uint32_t> order;  // tri_order_: leaves own contiguous ranges of it
    size_t non_manifold = 0;
    size_t tris() const { return idx.size() / 3; }
};

inline uint64_t mesh_edge_key(uint32_t a, uint32_t b) {  // triangle_mesh.h:25-29
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return (uint64_t(lo) << 32) | uint64_t(hi);
}
inline uint64_t mesh_cell_key(int32_t x, int32_t y, int32_t z) {  // triangle_mesh.h:31-34
    auto u = [](int32_t q) { return uint64_t(uint32_t(q)) & 0x1FFFFFu; };
    return (u(x) << 42) | (u(y) << 21) | u(z);
}

// mesh_bvh.h:91-111: node bounds are the merge of the range's triangle bounds; a range of at most
// 4 triangles is a leaf; otherwise split at the median centroid on the longest extent's axis
// (std::nth_element, so the split is libstdc++'s introselect on the same comparator).
inline int32_t mesh_bvh_node(Mesh& m, size_t start, size_t end, const std::vector<AABB<float>>& tb,
                             const std::vector<V3<float>>& cen) {
    AABB<float> b;
    for (size_t i = start; i < end; ++i) b = AABB<float>::merge(b, tb[m.order[i]]);
    const size_t count = end - start;
    const int32_t id = int32_t(m.nodes.size());
    m.nodes.push_back(MeshNode{});
    m.nodes[size_t(id)].b = b;
    if (count <= 4) {
        m.nodes[size_t(id)].start = uint32_t(start);
        m.nodes[size_t(id)].count = uint32_t(count);
        return id;
    }
    const V3<float> ext = b.mx - b.mn;
    int axis = 0;
    if (ext.y > ext[axis]) axis = 1;
    if (ext.z > ext[axis]) axis = 2;
    const size_t mid = start + count / 2;
    std::nth_element(m.order.begin() + long(start), m.order.begin() + long(mid), m.order.begin() + long(end),
                     [&](uint32_t a, uint32_t c) { return cen[a][axis] < cen[c][axis]; });
    const int32_t l = mesh_bvh_node(m, start, mid, tb, cen);
    const int32_t r = mesh_bvh_node(m, mid, end, tb, cen);
    m.nodes[size_t(id)].left = l;
    m.nodes[size_t(id)].right = r;
    return id;
}
inline void mesh_bvh_build(Mesh& m) {  // mesh_bvh.h:31-51
    m.nodes.clear();
    const size_t nt = m.tris();
    m.order.resize(nt);
    if (!nt) return;
    std::vector<AABB<float>> tb(nt);
    std::vector<V3<float>> cen(nt);
    for (size_t t = 0; t < nt; ++t) {
        m.order[t] = uint32_t(t);
        const V3<float>& a = m.v[m.idx[3 * t]];
        const V3<float>& b = m.v[m.idx[3 * t + 1]];
        const V3<float>& c = m.v[m.idx[3 * t + 2]];
        tb[t].mn = vmin(a, vmin(b, c));
        tb[t].mx = vmax(a, vmax(b, c));
        cen[t] = (a + b + c) * (1.0f / 3.0f);
    }
    mesh_bvh_node(m, 0, nt, tb, cen);
}
// mesh_bvh.h:122-132: every triangle of every leaf whose bounds (and so every ancestor's) overlap
template <class T, class Fn> void mesh_bvh_query(const Mesh& m, const AABB<T>& q, Fn&& fn) {
    if (m.nodes.empty()) return;
    std::vector<int32_t> stack{0};
    while (!stack.empty()) {
        const MeshNode& nd = m.nodes[size_t(stack.back())];
        stack.pop_back();
        AABB<T> b;
        b.mn = V3<T>(T(nd.b.mn.x), T(nd.b.mn.y), T(nd.b.mn.z));
        b.mx = V3<T>(T(nd.b.mx.x), T(nd.b.mx.y), T(nd.b.mx.z));
        if (!b.overlaps(q)) continue;
        if (nd.left < 0) {
            for (uint32_t i = 0; i < nd.count; ++i) fn(m.order[nd.start + i]);
        } else {
            stack.push_back(nd.right);
            stack.push_back(nd.left);
        }
    }
}

// TriangleMesh::from_triangles (triangle_mesh.h:75-170). Indices must be < pos.size().
inline Mesh mesh_from_triangles(const std::vector<V3<float>>& pos, const std::vector<uint32_t>& ind) {
    Mesh m;
    if (pos.empty() || ind.size() < 3) return m;
    V3<float> bmn = pos[0], bmx = pos[0];
    for (const V3<float>& p : pos) {
        bmn = vmin(bmn, p);
        bmx = vmax(bmx, p);
    }
    const float diagonal = length(bmx - bmn);
    const float eps = diagonal > 0.0f ? diagonal * 1e-4f : 1e-5f;
    const float cell = eps > 0.0f ? eps : 1e-5f;
    // weld: each position joins the first welded vertex within eps found over the 27 cells around
    // its own (dx, dy, dz ascending; within a cell in the hash multimap's order), else starts one
    std::vector<uint32_t> remap(pos.size());
    std::unordered_multimap<uint64_t, uint32_t> grid;
    grid.reserve(pos.size());
    for (size_t i = 0; i < pos.size(); ++i) {
        const V3<float>& p = pos[i];
        const int32_t cx = int32_t(std::floor(p.x / cell)), cy = int32_t(std::floor(p.y / cell)),
                      cz = int32_t(std::floor(p.z / cell));
        uint32_t found = UINT32_MAX;
        for (int dx = -1; dx <= 1 && found == UINT32_MAX; ++dx)
            for (int dy = -1; dy <= 1 && found == UINT32_MAX; ++dy)
                for (int dz = -1; dz <= 1 && found == UINT32_MAX; ++dz) {
                    auto range = grid.equal_range(mesh_cell_key(cx + dx, cy + dy, cz + dz));
                    for (auto it = range.first; it != range.second; ++it)
                        if (length(m.v[it->second] - p) <= eps) {
                            found = it->second;
                            break;
                        }
                }
        if (found == UINT32_MAX) {
            found = uint32_t(m.v.size());
            m.v.push_back(p);
            grid.emplace(mesh_cell_key(cx, cy, cz), found);
        }
        remap[i] = found;
    }
    // triangles through the weld, dropping those it collapsed or left without area
    for (size_t t = 0; t + 2 < ind.size(); t += 3) {
        const uint32_t a = remap[ind[t]], b = remap[ind[t + 1]], c = remap[ind[t + 2]];
        if (a == b || b == c || a == c) continue;
        if (length(cross(m.v[b] - m.v[a], m.v[c] - m.v[a])) <= 1e-8f) continue;
        m.idx.push_back(a);
        m.idx.push_back(b);
        m.idx.push_back(c);
    }
    const size_t nt = m.tris();
    m.n.resize(nt);
    for (size_t t = 0; t < nt; ++t) {
        const V3<float>& a = m.v[m.idx[3 * t]];
        const V3<float>& b = m.v[m.idx[3 * t + 1]];
        const V3<float>& c = m.v[m.idx[3 * t + 2]];
        m.n[t] = normalize(cross(b - a, c - a));
    }
    // an edge owned by exactly two triangles links them; by more, it is counted as non-manifold
    m.adj.assign(nt, {k_no_neighbor, k_no_neighbor, k_no_neighbor});
    std::unordered_map<uint64_t, std::vector<std::pair<uint32_t, uint32_t>>> owners;
    for (uint32_t t = 0; t < nt; ++t)
        for (uint32_t e = 0; e < 3; ++e)
            owners[mesh_edge_key(m.idx[3 * t + e], m.idx[3 * t + (e + 1) % 3])].push_back({t, e});
    for (const auto& kv : owners) {
        const auto& o = kv.second;
        if (o.size() == 2) {
            m.adj[o[0].first][o[0].second] = o[1].first;
            m.adj[o[1].first][o[1].second] = o[0].first;
        } else if (o.size() > 2) {
            ++m.non_manifold;
        }
    }
    for (const V3<float>& p : m.v) {  // triangle_mesh.h:290-297
        m.bounds.mn = vmin(m.bounds.mn, p);
        m.bounds.mx = vmax(m.bounds.mx, p);
    }
    mesh_bvh_build(m);
    return m;
}

// ---- transforms: glm 0.9.9 mat4 ----

template <class T> struct M4 {  // column-major: m[col][row]
    T m[4][4] = {};
};
// mesh_contact.h:303-307 (and shape.h:216-220): mat4(mat3_cast(rot * local_rotation) * scale) with
// column 3 = (pos + rot * local_center, 1)
template <class T> M4<T> mesh_world_transform(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot) {
    const M3<T> r = mat3_cast(qmul(rot, s.lr));
    M4<T> o;
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) o.m[c][k] = r.c[c][k] * s.mesh_scale;
    const V3<T> t = pos + qrot(rot, s.lc);
    o.m[3][0] = t.x; o.m[3][1] = t.y; o.m[3][2] = t.z; o.m[3][3] = T(1);
    return o;
}
// vec3(m * vec4(p, 1)): glm's (m0*x + m1*y) + (m2*z + m3*w)
template <class T> inline V3<T> m4_point(const M4<T>& m, const V3<T>& p) {
    V3<T> o;
    for (int k = 0; k < 3; ++k) o[k] = (m.m[0][k] * p.x + m.m[1][k] * p.y) + (m.m[2][k] * p.z + m.m[3][k] * T(1));
    return o;
}
// glm::inverse of a mat4 (func_matrix.inl compute_inverse<4,4>): 2x2 cofactors Coef00..Coef23,
// Fac0..Fac5, Inv0..Inv3 with the alternating signs, scaled by 1 / (m[0] . first row)
template <class T> M4<T> m4_inverse(const M4<T>& M) {
    auto m = [&](int c, int r) { return M.m[c][r]; };
    const T c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3), c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3),
            c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
    const T c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3), c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3),
            c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
    const T c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2), c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2),
            c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
    const T c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3), c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3),
            c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
    const T c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2), c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2),
            c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
    const T c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1), c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1),
            c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
    const T f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
    const T f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    const T v0[4] = {m(1, 0), m(0, 0), m(0, 0), m(0, 0)}, v1[4] = {m(1, 1), m(0, 1), m(0, 1), m(0, 1)};
    const T v2[4] = {m(1, 2), m(0, 2), m(0, 2), m(0, 2)}, v3[4] = {m(1, 3), m(0, 3), m(0, 3), m(0, 3)};
    const T sa[4] = {T(1), T(-1), T(1), T(-1)}, sb[4] = {T(-1), T(1), T(-1), T(1)};
    M4<T> inv;
    for (int i = 0; i < 4; ++i) {
        inv.m[0][i] = ((v1[i] * f0[i] - v2[i] * f1[i]) + v3[i] * f2[i]) * sa[i];
        inv.m[1][i] = ((v0[i] * f0[i] - v2[i] * f3[i]) + v3[i] * f4[i]) * sb[i];
        inv.m[2][i] = ((v0[i] * f1[i] - v1[i] * f3[i]) + v3[i] * f5[i]) * sa[i];
        inv.m[3][i] = ((v0[i] * f2[i] - v1[i] * f4[i]) + v2[i] * f5[i]) * sb[i];
    }
    const T d0 = m(0, 0) * inv.m[0][0], d1 = m(0, 1) * inv.m[1][0], d2 = m(0, 2) * inv.m[2][0], d3 = m(0, 3) * inv.m[3][0];
    const T det = (d0 + d1) + (d2 + d3);
    const T one_over = T(1) / det;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) inv.m[c][r] = inv.m[c][r] * one_over;
    return inv;
}
// AABB::transform (aabb.h:83-95): the 8 corners through m, refitted
template <class T> AABB<T> aabb_transform(const AABB<T>& b, const M4<T>& m) {
    AABB<T> out;
    for (int i = 0; i < 8; ++i) {
        const V3<T> c((i & 1) ? b.mx.x : b.mn.x, (i & 2) ? b.mx.y : b.mn.y, (i & 4) ? b.mx.z : b.mn.z);
        const V3<T> w = m4_point(m, c);
        out.mn = vmin(out.mn, w);
        out.mx = vmax(out.mx, w);
    }
    return out;
}
template <class T> inline V3<T> mv(const V3<float>& p) { return V3<T>(T(p.x), T(p.y), T(p.z)); }

template <class T> AABB<T> mesh_world_bounds(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot) {  // shape.h:216-220
    const Mesh* mesh = static_cast<const Mesh*>(s.mesh);
    if (!mesh) return AABB<T>{};
    AABB<T> lb;
    lb.mn = mv<T>(mesh->bounds.mn);
    lb.mx = mv<T>(mesh->bounds.mx);
    return aabb_transform(lb, mesh_world_transform(s, pos, rot));
}

// ---- per-triangle geometry (mesh_contact.h) ----

// closest_point_on_triangle (mesh_contact.h:61-98): Ericson 5.1.5's Voronoi-region cases
template <class T> V3<T> closest_on_triangle(const V3<T>& p, const V3<T>& a, const V3<T>& b, const V3<T>& c) {
    const V3<T> ab = b - a, ac = c - a, ap = p - a;
    const T d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= T(0) && d2 <= T(0)) return a;
    const V3<T> bp = p - b;
    const T d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= T(0) && d4 <= d3) return b;
    const T vc = d1 * d4 - d3 * d2;
    if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
        const T v = d1 / (d1 - d3);
        return a + v * ab;
    }
    const V3<T> cp = p - c;
    const T d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= T(0) && d5 <= d6) return c;
    const T vb = d5 * d2 - d1 * d6;
    if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
        const T w = d2 / (d2 - d6);
        return a + w * ac;
    }
    const T va = d3 * d6 - d5 * d4;
    if (va <= T(0) && (d4 - d3) >= T(0) && (d5 - d6) >= T(0)) {
        const T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return b + w * (c - b);
    }
    const T denom = T(1) / (va + vb + vc);
    const T v = vb * denom, w = vc * denom;
    return a + ab * v + ac * w;
}
// barycentric_ (mesh_contact.h:247-258)
template <class T> V3<T> barycentric(const V3<T>& p, const V3<T>& a, const V3<T>& b, const V3<T>& c) {
    const V3<T> v0 = b - a, v1 = c - a, v2 = p - a;
    const T d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    const T denom = d00 * d11 - d01 * d01;
    if (std::abs(denom) < k_epsilon<T>()) return V3<T>(T(1), T(0), T(0));
    const T v = (d11 * d20 - d01 * d21) / denom;
    const T w = (d00 * d21 - d01 * d20) / denom;
    return V3<T>(T(1) - v - w, v, w);
}
template <class T> inline bool bary_outside(const V3<T>& b) { return b.x < T(-1e-4f) || b.y < T(-1e-4f) || b.z < T(-1e-4f); }
// classify_triangle_edge_ (mesh_contact.h:265-272)
template <class T> int classify_edge(const V3<T>& b) {
    const T eps = T(1e-3f);
    const bool na = b.x < eps, nb = b.y < eps, nc = b.z < eps;
    if (nc && !na && !nb) return 0;
    if (na && !nb && !nc) return 1;
    if (nb && !na && !nc) return 2;
    return -1;
}
// correct_mesh_normal_ (mesh_contact.h:291-301)
template <class T> V3<T> correct_normal(int edge, const V3<T>& fn, const V3<T>& nn, bool has, const V3<T>& raw) {
    if (edge < 0) return dot(raw, fn) > T(0) ? fn : raw;
    if (!has) return raw;
    if (dot(fn, nn) > T(0.999f)) return fn;
    return dot(raw, fn) >= dot(raw, nn) ? fn : nn;
}

// One triangle posed in the world: vertices through the mesh transform, rotated unit normal.
template <class T> struct WTri {
    V3<T> v[3];
    V3<T> n;
};
template <class T> WTri<T> world_triangle(const Mesh& mesh, const M4<T>& mt, const Q<T>& mwr, uint32_t tri) {
    WTri<T> w;
    for (int k = 0; k < 3; ++k) w.v[k] = m4_point(mt, mv<T>(mesh.v[mesh.idx[3 * tri + uint32_t(k)]]));
    w.n = normalize(qrot(mwr, mv<T>(mesh.n[tri])));
    return w;
}

// mesh_contact_candidates (mesh_contact.h:330-340): the BVH's triangles for the shape's bounds
// (a box's grown by 1.5 x its half diagonal, a hull's by 1.5 x half its local-bounds diagonal x
// scale) mapped into the mesh's local frame, ascending
template <class T> void mesh_candidates(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot, const Mesh& mesh,
                                        const M4<T>& mt, std::vector<uint32_t>& out) {
    out.clear();
    AABB<T> probe = world_bounds(s, pos, rot);
    if (s.type == PXG_BOX) {
        probe = probe.expand(T(1.5f) * length(s.he));
    } else if (s.type == PXG_CONVEX_HULL && s.hull) {
        const Hull<T>* h = static_cast<const Hull<T>*>(s.hull);
        probe = probe.expand(T(1.5f) * T(0.5f) * length(h->lb.mx - h->lb.mn) * s.hull_scale);
    }
    mesh_bvh_query(mesh, aabb_transform(probe, m4_inverse(mt)), [&](uint32_t t) { out.push_back(t); });
    std::sort(out.begin(), out.end());
}

template <class T> struct MeshCand {  // generate_box_mesh_contacts_'s Candidate
    V3<T> point, normal, key;
    T pen = 0;
    uint32_t fid = 0;
};

// The shared tail of generate_box_mesh_contacts_ and generate_hull_mesh_contacts_
// (mesh_contact.h:565-614 / 771-820): recovery when nothing else was found, else the deepest
// point, the one farthest from it, the one farthest from their line, the one farthest from all
// three (each must beat the best so far by 1e-6 and score above 1e-6), in candidate order, under
// the penetration-weighted mean normal.
template <class T> bool mesh_reduce(const std::vector<MeshCand<T>>& all, bool recovery_valid, const MeshCand<T>& recovery,
                                    Manifold<T>& out) {
    if (all.empty()) {
        if (!recovery_valid) return false;
        out.add_point(recovery.point, recovery.pen, recovery.fid);
        out.n = recovery.normal;
        return true;
    }
    size_t pick[4] = {0, 0, 0, 0};
    size_t picked = 0;
    auto taken = [&](size_t i) { return std::find(pick, pick + picked, i) != pick + picked; };
    auto best = [&](auto&& score) {
        size_t arg = all.size();
        T top = T(-1);
        for (size_t i = 0; i < all.size(); ++i) {
            if (taken(i)) continue;
            const T sc = score(all[i]);
            if (sc > top + T(1e-6f)) { top = sc; arg = i; }
        }
        if (arg < all.size() && top > T(1e-6f)) pick[picked++] = arg;
    };
    size_t deepest = 0;
    for (size_t i = 1; i < all.size(); ++i)
        if (all[i].pen > all[deepest].pen + T(1e-6f)) deepest = i;
    pick[picked++] = deepest;
    const V3<T> p0 = all[pick[0]].point;
    best([&](const MeshCand<T>& c) { return length(c.point - p0); });
    if (picked == 2) {
        const V3<T> p1 = all[pick[1]].point;
        best([&](const MeshCand<T>& c) { return length(cross(p1 - p0, c.point - p0)); });
    }
    if (picked == 3) {
        const V3<T> p1 = all[pick[1]].point, p2 = all[pick[2]].point;
        best([&](const MeshCand<T>& c) {
            T a = length(c.point - p0), b = length(c.point - p1), d = length(c.point - p2);
            T r = a;  // std::min({a, b, d}): the first smallest
            if (b < r) r = b;
            if (d < r) r = d;
            return r;
        });
    }
    std::sort(pick, pick + picked);
    V3<T> normal(T(0));
    for (size_t i = 0; i < picked; ++i) {
        const MeshCand<T>& c = all[pick[i]];
        out.add_point(c.point, c.pen, c.fid);
        normal = normal + c.normal * gmax(c.pen, T(1e-4f));
    }
    const T len = length(normal);
    out.n = len > T(1e-6f) ? normal / len : all[pick[0]].normal;
    return true;
}

// generate_box_mesh_contacts_ (mesh_contact.h:471-615)
template <class T> bool box_mesh_contacts(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot, const Mesh& mesh,
                                          const M4<T>& mt, const Q<T>& mwr, const std::vector<uint32_t>& tris,
                                          Manifold<T>& out) {
    const OBB<T> box = world_obb(s, pos, rot);
    V3<T> ax[3];
    obb_axes(box, ax);
    const V3<T>& he = box.he;
    V3<T> corner[8];
    for (uint32_t c = 0; c < 8; ++c)
        corner[c] = box.c + ((c & 1u) ? he.x : -he.x) * ax[0] + ((c & 2u) ? he.y : -he.y) * ax[1] +
                    ((c & 4u) ? he.z : -he.z) * ax[2];
    const T inside_eps = T(1e-4f) * gmax(T(1), length(he));
    MeshCand<T> corners[8];
    bool valid[8] = {};
    std::vector<MeshCand<T>> verts;
    MeshCand<T> recovery;
    bool recovery_valid = false;
    for (uint32_t tri : tris) {
        const WTri<T> w = world_triangle(mesh, mt, mwr, tri);
        const V3<T>& n = w.n;
        const T center_dist = dot(box.c - w.v[0], n);
        if (center_dist <= T(0)) {
            const V3<T> under = box.c - center_dist * n;
            if (bary_outside(barycentric(under, w.v[0], w.v[1], w.v[2]))) continue;
            const T support = std::abs(dot(he.x * ax[0], n)) + std::abs(dot(he.y * ax[1], n)) + std::abs(dot(he.z * ax[2], n));
            const T exit = support - center_dist;
            if (recovery_valid && recovery.pen <= exit) continue;
            recovery_valid = true;
            recovery = MeshCand<T>{under, n, under, exit, 0x80000000u | tri};
            continue;
        }
        T depth = T(0);
        T dist[8];
        for (uint32_t c = 0; c < 8; ++c) {
            dist[c] = dot(corner[c] - w.v[0], n);
            depth = gmax(depth, -dist[c]);
        }
        if (depth <= T(0)) continue;
        for (uint32_t c = 0; c < 8; ++c) {
            if (dist[c] >= T(0)) continue;
            const V3<T> plane_point = corner[c] - dist[c] * n;
            if (bary_outside(barycentric(plane_point, w.v[0], w.v[1], w.v[2]))) continue;
            const T pen = -dist[c];
            if (valid[c] && corners[c].pen <= pen) continue;
            valid[c] = true;
            corners[c] = MeshCand<T>{corner[c] + n * (pen * T(0.5f)), n, corner[c], pen, c};
        }
        for (uint32_t k = 0; k < 3; ++k) {
            const V3<T> d = w.v[k] - box.c;
            if (std::abs(dot(d, ax[0])) > he.x + inside_eps || std::abs(dot(d, ax[1])) > he.y + inside_eps ||
                std::abs(dot(d, ax[2])) > he.z + inside_eps)
                continue;
            bool seen = false;
            for (const MeshCand<T>& c : verts) seen = seen || length(c.key - w.v[k]) < inside_eps;
            if (seen) continue;
            verts.push_back(MeshCand<T>{w.v[k] - n * (depth * T(0.5f)), n, w.v[k], depth, 8u + tri * 3u + k});
        }
    }
    std::vector<MeshCand<T>> all;
    for (uint32_t c = 0; c < 8; ++c)
        if (valid[c]) all.push_back(corners[c]);
    all.insert(all.end(), verts.begin(), verts.end());
    return mesh_reduce(all, recovery_valid, recovery, out);
}

// generate_hull_mesh_contacts_ (mesh_contact.h:645-821)
template <class T> bool hull_mesh_contacts(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot, const Mesh& mesh,
                                           const M4<T>& mt, const Q<T>& mwr, const std::vector<uint32_t>& tris,
                                           Manifold<T>& out) {
    if (!s.hull) return false;
    const PosedHull<T> ph = world_posed_hull(s, pos, rot);
    const Hull<T>& hull = *ph.h;
    const std::vector<V3<T>> hv = hull_world_vertices(ph);
    const V3<T> hc = ph.pos + qrot(ph.rot, hull.centroid * ph.scale);
    const T inside_eps = T(1e-4f) * gmax(T(1), length(hull.lb.mx - hull.lb.mn) * ph.scale);
    std::vector<MeshCand<T>> corners(hv.size());
    std::vector<bool> valid(hv.size(), false);
    std::vector<MeshCand<T>> verts;
    MeshCand<T> recovery;
    bool recovery_valid = false;
    std::vector<T> dist(hv.size());
    for (uint32_t tri : tris) {
        const WTri<T> w = world_triangle(mesh, mt, mwr, tri);
        const V3<T>& n = w.n;
        const T center_dist = dot(hc - w.v[0], n);
        if (center_dist <= T(0)) {
            const V3<T> under = hc - center_dist * n;
            if (bary_outside(barycentric(under, w.v[0], w.v[1], w.v[2]))) continue;
            T support = T(0);
            for (const V3<T>& p : hv) support = gmax(support, dot(p - hc, n));
            const T exit = support - center_dist;
            if (recovery_valid && recovery.pen <= exit) continue;
            recovery_valid = true;
            recovery = MeshCand<T>{under, n, under, exit, 0x80000000u | tri};
            continue;
        }
        T depth = T(0);
        for (size_t c = 0; c < hv.size(); ++c) {
            dist[c] = dot(hv[c] - w.v[0], n);
            depth = gmax(depth, -dist[c]);
        }
        if (depth <= T(0)) continue;
        for (size_t c = 0; c < hv.size(); ++c) {
            if (dist[c] >= T(0)) continue;
            const V3<T> plane_point = hv[c] - dist[c] * n;
            if (bary_outside(barycentric(plane_point, w.v[0], w.v[1], w.v[2]))) continue;
            const T pen = -dist[c];
            if (valid[c] && corners[c].pen <= pen) continue;
            valid[c] = true;
            corners[c] = MeshCand<T>{hv[c] + n * (pen * T(0.5f)), n, hv[c], pen, uint32_t(c)};
        }
        for (uint32_t k = 0; k < 3; ++k) {
            bool inside = true;
            for (size_t f = 0; f < hull.faces() && inside; ++f) {
                const V3<T> fn = ph.fnormal(f);
                const T fd = dot(fn, ph.vtx(hull.fv[hull.fs[f]]));
                if (dot(fn, w.v[k]) - fd > inside_eps) inside = false;
            }
            if (!inside) continue;
            bool seen = false;
            for (const MeshCand<T>& c : verts) seen = seen || length(c.key - w.v[k]) < inside_eps;
            if (seen) continue;
            verts.push_back(MeshCand<T>{w.v[k] - n * (depth * T(0.5f)), n, w.v[k], depth, uint32_t(hv.size()) + tri * 3u + k});
        }
    }
    std::vector<MeshCand<T>> all;
    for (size_t c = 0; c < hv.size(); ++c)
        if (valid[c]) all.push_back(corners[c]);
    all.insert(all.end(), verts.begin(), verts.end());
    return mesh_reduce(all, recovery_valid, recovery, out);
}

// mesh_triangle_contacts' sphere/capsule branch (mesh_contact.h:351-465, with
// sphere_or_capsule_vs_triangle_contact_ at 110-131): emit(point, normal, pen, fid)
template <class T, class Emit>
void round_triangle_contact(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot, const Mesh& mesh, const M4<T>& mt,
                            const Q<T>& mwr, uint32_t tri, Emit&& emit) {
    const WTri<T> w = world_triangle(mesh, mt, mwr, tri);
    const V3<T> &v0 = w.v[0], &v1 = w.v[1], &v2 = w.v[2], &fn = w.n;
    V3<T> on_shape, on_tri;
    if (s.type == PXG_SPHERE) {
        on_shape = world_sphere(s, pos, rot).c;
        on_tri = closest_on_triangle(on_shape, v0, v1, v2);
    } else {
        const Cap<T> cap = world_capsule(s, pos, rot);
        on_tri = (v0 + v1 + v2) * T(1.0f / 3.0f);
        for (int i = 0; i < 6; ++i) {
            on_shape = closest_point_on_segment(on_tri, cap.a, cap.b);
            on_tri = closest_on_triangle(on_shape, v0, v1, v2);
        }
    }
    const T radius = s.type == PXG_SPHERE ? s.radius : s.cr;
    const T pen = radius - length(on_shape - on_tri);
    const bool touching = pen > T(0);
    const T core_side = dot(on_shape - v0, fn);
    if (core_side < T(0)) {  // embedded: out along the face, the whole way, if under this triangle
        const V3<T> under = on_shape - core_side * fn;
        if (bary_outside(barycentric(under, v0, v1, v2))) return;
        const T exit = radius - core_side;
        emit(under + fn * (exit * T(0.5f)), fn, exit, tri);
        return;
    }
    if (!touching) return;
    const V3<T> diff = on_shape - on_tri;
    const T dist = length(diff);
    const V3<T> raw = dist > k_epsilon<T>() ? diff / dist : fn;
    const int edge = classify_edge(barycentric(on_tri, v0, v1, v2));
    V3<T> nn(T(0));
    bool has = false;
    if (edge >= 0) {
        const uint32_t nt = mesh.adj[tri][size_t(edge)];
        if (nt != k_no_neighbor) {
            has = true;
            nn = normalize(qrot(mwr, mv<T>(mesh.n[nt])));
        }
    }
    const V3<T> normal = correct_normal(edge, fn, nn, has, raw);
    emit(on_tri + normal * (pen * T(0.5f)), normal, pen, tri);
}

// generate_mesh_contacts (mesh_contact.h:838-884): the manifold between a convex shape and a
// mesh shape, normal from the mesh toward the shape. use_bvh false scans every triangle.
template <class T> bool mesh_contacts(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot, const Shape<T>& ms,
                                      const V3<T>& mpos, const Q<T>& mrot, Manifold<T>& out, bool use_bvh = true) {
    out = Manifold<T>{};
    const Mesh* mesh = static_cast<const Mesh*>(ms.mesh);
    if (!mesh) return false;
    const M4<T> mt = mesh_world_transform(ms, mpos, mrot);
    const Q<T> mwr = qmul(mrot, ms.lr);
    std::vector<uint32_t> tris;
    if (use_bvh) {
        mesh_candidates(s, pos, rot, *mesh, mt, tris);
    } else {
        tris.resize(mesh->tris());
        for (size_t t = 0; t < tris.size(); ++t) tris[t] = uint32_t(t);
    }
    if (s.type == PXG_BOX) return box_mesh_contacts(s, pos, rot, *mesh, mt, mwr, tris, out);
    if (s.type == PXG_CONVEX_HULL) return hull_mesh_contacts(s, pos, rot, *mesh, mt, mwr, tris, out);
    bool any = false;
    V3<T> best_normal(T(0), T(0), T(1));
    auto emit = [&](const V3<T>& p, const V3<T>& n, T pen, uint32_t fid) {
        if (out.count >= 4) return;
        out.add_point(p, pen, fid);
        best_normal = n;
        any = true;
    };
    for (uint32_t tri : tris) {
        if (out.count >= 4) break;
        round_triangle_contact(s, pos, rot, *mesh, mt, mwr, tri, emit);
    }
    if (!any) return false;
    out.n = best_normal;
    return true;
}

// narrowphase.h:313-320: the mesh branches of generate_contacts (no speculative contacts, no
// mesh against mesh); false when neither shape is a mesh
template <class T> bool mesh_pair_contacts(const Shape<T>& sa, const V3<T>& pa, const Q<T>& ra, const Shape<T>& sb,
                                           const V3<T>& pb, const Q<T>& rb, Manifold<T>& out, bool& hit) {
    const bool ma = sa.type == PXG_TRIANGLE_MESH, mb = sb.type == PXG_TRIANGLE_MESH;
    if (!ma && !mb) return false;
    hit = false;
    if (ma && mb) return true;
    if (mb) {
        hit = mesh_contacts(sa, pa, ra, sb, pb, rb, out);
        if (hit) out.n = -out.n;
    } else {
        hit = mesh_contacts(sb, pb, rb, sa, pa, ra, out);
    }
    return true;
}

}  // namespace pxo
