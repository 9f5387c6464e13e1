// Convex-against-triangle-mesh narrowphase on the GPU (collision/mesh_contact.h): a sphere,
// capsule, box or convex hull against a static or kinematic mesh, with the reference's candidate
// triangles, per-triangle tests, internal-edge correction, embedded recovery, 4-point spread
// reduction and feature ids. One thread per pair.
//
// Candidate triangles: the reference queries its triangle tree with the shape's world bounds
// (grown for a box or hull) mapped into the mesh frame through glm::inverse of the mesh
// transform, then visits the hits in ascending index order. Here the same tree (cooked on the
// host) is walked with a stack; each walk keeps the kMeshWindow smallest hits above the last
// index visited, so the triangles are visited in the same ascending order, a window at a time,
// however many there are.
//
// A box or hull keeps its corner candidates (one per corner, the least penetrating triangle's)
// and the mesh vertices found inside it (at most kMeshVerts; more raise ERR_MESH_VERTICES).
#pragma once
#include "pxg_hull.h"

namespace pxd {

constexpr int SH_MESH = 4;                // pxg_shape_type PXG_TRIANGLE_MESH
constexpr uint32_t kMeshNone = 0xFFFFFFFFu;
constexpr int kMeshWindow = 32;           // candidate triangles per walk of the tree
constexpr int kMeshStack = 64;            // walk stack; the cooked tree is at most kMeshStack - 2 levels deep
constexpr int kMeshVerts = 32;            // mesh vertices inside one box or hull

struct MeshTab {
    const uint4* meta;      // per mesh: vertex offset, triangle offset, node offset, triangle count
    const float4* lb;       // per mesh: bounds of the welded vertices (min, max)
    const float4* v;        // welded vertices
    const uint4* tri;       // per triangle: mesh-local vertex indices
    const float4* n;        // per triangle: unit face normal
    const uint4* adj;       // per triangle: neighbour across v0-v1, v1-v2, v2-v0 (mesh-local) or kMeshNone
    const float4* node_lo;  // per node: bounds min, left child as int bits (-1 for a leaf)
    const float4* node_hi;  // per node: bounds max, right child as int bits
    const uint2* node_tri;  // per node: a leaf's first entry in order and triangle count
    const uint32_t* order;  // triangle order (mesh-local triangle indices), per mesh from its triangle offset
};

struct m4 { float m[4][4]; };  // glm mat4, column-major: m[col][row]

// mesh_world_transform (mesh_contact.h:303-307): mat4(mat3_cast(rot * local_rotation) * scale),
// column 3 = (pos + rot * local_center, 1); the mesh scale is in the shape's radius
PXD m4 mesh_transform(const ShapeGeom& s, v3 pos, quat rot) {
    const m3 r = mat3_cast(qmul(rot, s.lr));
    const v3 t = pos + qrot(rot, s.lc);
    m4 o;
    const v3 c[3] = {r.c0, r.c1, r.c2};
    for (int k = 0; k < 3; ++k) {
        o.m[k][0] = c[k].x * s.radius;
        o.m[k][1] = c[k].y * s.radius;
        o.m[k][2] = c[k].z * s.radius;
        o.m[k][3] = 0.0f;
    }
    o.m[3][0] = t.x; o.m[3][1] = t.y; o.m[3][2] = t.z; o.m[3][3] = 1.0f;
    return o;
}
// vec3(m * vec4(p, 1)): (m0*x + m1*y) + (m2*z + m3*1)
PXD v3 m4_point(const m4& m, v3 p) {
    return mk((m.m[0][0] * p.x + m.m[1][0] * p.y) + (m.m[2][0] * p.z + m.m[3][0] * 1.0f),
              (m.m[0][1] * p.x + m.m[1][1] * p.y) + (m.m[2][1] * p.z + m.m[3][1] * 1.0f),
              (m.m[0][2] * p.x + m.m[1][2] * p.y) + (m.m[2][2] * p.z + m.m[3][2] * 1.0f));
}
// glm::inverse (func_matrix.inl compute_inverse<4,4>): the 2x2 cofactors of the lower rows, the
// signed 3x3 minors, scaled by 1 / (first column . first row of the adjugate)
PXD m4 m4_inverse(const m4& M) {
    const float(*m)[4] = M.m;
    const float c00 = m[2][2] * m[3][3] - m[3][2] * m[2][3], c02 = m[1][2] * m[3][3] - m[3][2] * m[1][3],
                c03 = m[1][2] * m[2][3] - m[2][2] * m[1][3];
    const float c04 = m[2][1] * m[3][3] - m[3][1] * m[2][3], c06 = m[1][1] * m[3][3] - m[3][1] * m[1][3],
                c07 = m[1][1] * m[2][3] - m[2][1] * m[1][3];
    const float c08 = m[2][1] * m[3][2] - m[3][1] * m[2][2], c10 = m[1][1] * m[3][2] - m[3][1] * m[1][2],
                c11 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
    const float c12 = m[2][0] * m[3][3] - m[3][0] * m[2][3], c14 = m[1][0] * m[3][3] - m[3][0] * m[1][3],
                c15 = m[1][0] * m[2][3] - m[2][0] * m[1][3];
    const float c16 = m[2][0] * m[3][2] - m[3][0] * m[2][2], c18 = m[1][0] * m[3][2] - m[3][0] * m[1][2],
                c19 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
    const float c20 = m[2][0] * m[3][1] - m[3][0] * m[2][1], c22 = m[1][0] * m[3][1] - m[3][0] * m[1][1],
                c23 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
    const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
    const float f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    const float a0[4] = {m[1][0], m[0][0], m[0][0], m[0][0]}, a1[4] = {m[1][1], m[0][1], m[0][1], m[0][1]};
    const float a2[4] = {m[1][2], m[0][2], m[0][2], m[0][2]}, a3[4] = {m[1][3], m[0][3], m[0][3], m[0][3]};
    const float sa[4] = {1.0f, -1.0f, 1.0f, -1.0f}, sb[4] = {-1.0f, 1.0f, -1.0f, 1.0f};
    m4 inv;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        inv.m[0][i] = ((a1[i] * f0[i] - a2[i] * f1[i]) + a3[i] * f2[i]) * sa[i];
        inv.m[1][i] = ((a0[i] * f0[i] - a2[i] * f3[i]) + a3[i] * f4[i]) * sb[i];
        inv.m[2][i] = ((a0[i] * f1[i] - a1[i] * f3[i]) + a3[i] * f5[i]) * sa[i];
        inv.m[3][i] = ((a0[i] * f2[i] - a1[i] * f4[i]) + a2[i] * f5[i]) * sb[i];
    }
    const float d0 = m[0][0] * inv.m[0][0], d1 = m[0][1] * inv.m[1][0], d2 = m[0][2] * inv.m[2][0],
                d3 = m[0][3] * inv.m[3][0];
    const float one_over = 1.0f / ((d0 + d1) + (d2 + d3));
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) inv.m[c][r] = inv.m[c][r] * one_over;
    return inv;
}
// AABB::transform (aabb.h:83-95)
PXD void box_through(const m4& m, v3 lo, v3 hi, v3& mn, v3& mx) {
    mn = splat(3.402823466e+38f);
    mx = splat(-3.402823466e+38f);
    for (int i = 0; i < 8; ++i) {
        const v3 w = m4_point(m, mk((i & 1) ? hi.x : lo.x, (i & 2) ? hi.y : lo.y, (i & 4) ? hi.z : lo.z));
        mn = vmin(mn, w);
        mx = vmax(mx, w);
    }
}
// shape.h:216-220: the mesh's bounds through its world transform
PXD void mesh_bounds(const MeshTab& t, const ShapeGeom& s, v3 pos, quat rot, v3& mn, v3& mx) {
    box_through(mesh_transform(s, pos, rot), ld3(t.lb[2 * s.hull]), ld3(t.lb[2 * s.hull + 1]), mn, mx);
}

// ---- per-triangle geometry ----

// closest_point_on_triangle (mesh_contact.h:61-98)
PXD v3 closest_on_triangle(v3 p, v3 a, v3 b, v3 c) {
    const v3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    const v3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return a + (d1 / (d1 - d3)) * ab;
    const v3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return a + (d2 / (d2 - d6)) * ac;
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
    const float denom = 1.0f / (va + vb + vc);
    return a + ab * (vb * denom) + ac * (vc * denom);
}
// barycentric_ (mesh_contact.h:247-258)
PXD v3 barycentric(v3 p, v3 a, v3 b, v3 c) {
    const v3 e0 = b - a, e1 = c - a, e2 = p - a;
    const float d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1), d20 = dot(e2, e0), d21 = dot(e2, e1);
    const float denom = d00 * d11 - d01 * d01;
    if (fabsf(denom) < 1e-6f) return mk(1.0f, 0.0f, 0.0f);
    const float v = (d11 * d20 - d01 * d21) / denom;
    const float w = (d00 * d21 - d01 * d20) / denom;
    return mk(1.0f - v - w, v, w);
}
PXD bool outside(v3 b) { return b.x < -1e-4f || b.y < -1e-4f || b.z < -1e-4f; }
// classify_triangle_edge_ (mesh_contact.h:265-272)
PXD int tri_edge(v3 b) {
    const bool na = b.x < 1e-3f, nb = b.y < 1e-3f, nc = b.z < 1e-3f;
    if (nc && !na && !nb) return 0;
    if (na && !nb && !nc) return 1;
    if (nb && !na && !nc) return 2;
    return -1;
}

struct MeshPose {  // one mesh shape posed for a pair
    uint4 meta;
    m4 mt;
    quat rot;  // body rotation * local rotation: turns the stored normals
};
struct WTri { v3 v[3]; v3 n; };
PXD WTri world_tri(const MeshTab& t, const MeshPose& mp, uint32_t tri) {
    const uint4 ix = t.tri[mp.meta.y + tri];
    WTri w;
    w.v[0] = m4_point(mp.mt, ld3(t.v[mp.meta.x + ix.x]));
    w.v[1] = m4_point(mp.mt, ld3(t.v[mp.meta.x + ix.y]));
    w.v[2] = m4_point(mp.mt, ld3(t.v[mp.meta.x + ix.z]));
    w.n = normalize(qrot(mp.rot, ld3(t.n[mp.meta.y + tri])));
    return w;
}

// One walk of the tree: the (at most kMeshWindow) smallest candidate triangles above `after`
// (every candidate when `first`), ascending, in win; `more` when candidates were left out.
PXD int mesh_window(const MeshTab& t, const uint4& meta, v3 qlo, v3 qhi, bool first, uint32_t after, uint32_t* win,
                    bool& more) {
    int n = 0;
    more = false;
    int32_t stack[kMeshStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int32_t k = stack[--sp];
        const float4 lo = t.node_lo[meta.z + uint32_t(k)], hi = t.node_hi[meta.z + uint32_t(k)];
        if (!(lo.x <= qhi.x && hi.x >= qlo.x && lo.y <= qhi.y && hi.y >= qlo.y && lo.z <= qhi.z && hi.z >= qlo.z)) continue;
        const int32_t left = __float_as_int(lo.w);
        if (left >= 0) {
            stack[sp++] = __float_as_int(hi.w);
            stack[sp++] = left;
            continue;
        }
        const uint2 r = t.node_tri[meta.z + uint32_t(k)];
        for (uint32_t i = 0; i < r.y; ++i) {
            const uint32_t tri = t.order[meta.y + r.x + i];
            if (!first && tri <= after) continue;
            if (n == kMeshWindow) {
                more = true;
                if (tri > win[n - 1]) continue;
                --n;  // the largest makes room
            }
            int j = n;
            while (j > 0 && win[j - 1] > tri) {
                win[j] = win[j - 1];
                --j;
            }
            win[j] = tri;
            ++n;
        }
    }
    return n;
}
// Depth-first walks of the tree, left subtree first (MeshBVH::query / raycast, mesh_bvh.h:122-145):
// leaf(tri) for every triangle of every leaf whose bounds overlap [qlo, qhi] / the ray's slab test
// passes, in the reference's visiting order.
template <class Leaf>
PXD void mesh_walk_box(const MeshTab& t, const uint4& meta, v3 qlo, v3 qhi, Leaf&& leaf) {
    int32_t stack[kMeshStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int32_t k = stack[--sp];
        const float4 lo = t.node_lo[meta.z + uint32_t(k)], hi = t.node_hi[meta.z + uint32_t(k)];
        if (!(lo.x <= qhi.x && hi.x >= qlo.x && lo.y <= qhi.y && hi.y >= qlo.y && lo.z <= qhi.z && hi.z >= qlo.z)) continue;
        const int32_t left = __float_as_int(lo.w);
        if (left >= 0) {
            stack[sp++] = __float_as_int(hi.w);
            stack[sp++] = left;
            continue;
        }
        const uint2 r = t.node_tri[meta.z + uint32_t(k)];
        for (uint32_t i = 0; i < r.y; ++i) leaf(t.order[meta.y + r.x + i]);
    }
}
// Ray::intersect (geometry/ray.h:31-42): slab test with 1e-9 added to every direction component
PXD bool ray_slab(v3 o, v3 dir, float maxd, v3 lo, v3 hi, float& t) {
    const v3 nd = dir + splat(1e-9f);
    const v3 inv = mk(1.0f / nd.x, 1.0f / nd.y, 1.0f / nd.z);
    const v3 t0 = mk((lo.x - o.x) * inv.x, (lo.y - o.y) * inv.y, (lo.z - o.z) * inv.z);
    const v3 t1 = mk((hi.x - o.x) * inv.x, (hi.y - o.y) * inv.y, (hi.z - o.z) * inv.z);
    const v3 a = vmin(t0, t1), b = vmax(t0, t1);
    float tn = a.x;  // std::max({...}) / std::min({...}): the first extreme
    if (tn < a.y) tn = a.y;
    if (tn < a.z) tn = a.z;
    if (tn < 0.0f) tn = 0.0f;
    float tf = b.x;
    if (b.y < tf) tf = b.y;
    if (b.z < tf) tf = b.z;
    if (maxd < tf) tf = maxd;
    if (tf < tn) return false;
    t = tn;
    return true;
}
template <class Leaf>
PXD void mesh_walk_ray(const MeshTab& t, const uint4& meta, v3 o, v3 dir, float maxd, Leaf&& leaf) {
    int32_t stack[kMeshStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int32_t k = stack[--sp];
        const float4 lo = t.node_lo[meta.z + uint32_t(k)], hi = t.node_hi[meta.z + uint32_t(k)];
        float th;
        if (!ray_slab(o, dir, maxd, ld3(lo), ld3(hi), th)) continue;
        const int32_t left = __float_as_int(lo.w);
        if (left >= 0) {
            stack[sp++] = __float_as_int(hi.w);
            stack[sp++] = left;
            continue;
        }
        const uint2 r = t.node_tri[meta.z + uint32_t(k)];
        for (uint32_t i = 0; i < r.y; ++i) leaf(t.order[meta.y + r.x + i]);
    }
}
// ray_vs_triangle (geometry/ray.h:205-226): Moller-Trumbore
PXD bool ray_triangle(v3 o, v3 dir, float maxd, v3 a, v3 b, v3 c, float& t) {
    const v3 e1 = b - a, e2 = c - a;
    const v3 pv = cross(dir, e2);
    const float det = dot(e1, pv);
    if (det > -1e-6f && det < 1e-6f) return false;
    const float inv = 1.0f / det;
    const v3 tv = o - a;
    const float u = dot(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    const v3 qv = cross(tv, e1);
    const float v = dot(dir, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    const float cand = dot(e2, qv) * inv;
    if (cand < 0.0f || cand > maxd) return false;
    t = cand;
    return true;
}

// Calls visit(tri) for the candidate triangles in ascending order until it returns false.
template <class Visit>
PXD void mesh_candidates(const MeshTab& t, const uint4& meta, v3 qlo, v3 qhi, Visit&& visit) {
    uint32_t win[kMeshWindow];
    bool first = true;
    uint32_t after = 0;
    for (;;) {
        bool more;
        const int n = mesh_window(t, meta, qlo, qhi, first, after, win, more);
        for (int i = 0; i < n; ++i)
            if (!visit(win[i])) return;
        if (!more || n == 0) return;
        first = false;
        after = win[n - 1];
    }
}

// ---- the convex shape's side ----

// sphere_or_capsule_vs_triangle_contact_ and mesh_triangle_contacts' sphere/capsule branch
// (mesh_contact.h:110-131, 421-464): emits through add_point, keeping the last normal
PXD void round_tri(const MeshTab& t, const MeshPose& mp, uint32_t tri, const ShapeGeom& s, v3 pos, quat rot, Man& out,
                   v3& last_n) {
    const WTri w = world_tri(t, mp, tri);
    const v3 v0 = w.v[0], v1 = w.v[1], v2 = w.v[2], fn = w.n;
    v3 on_shape, on_tri;
    if (s.type == SH_SPHERE) {
        on_shape = world_sphere(s, pos, rot).c;
        on_tri = closest_on_triangle(on_shape, v0, v1, v2);
    } else {
        const Cap cap = world_capsule(s, pos, rot);
        on_tri = (v0 + v1 + v2) * (1.0f / 3.0f);
        for (int i = 0; i < 6; ++i) {
            on_shape = closest_point_on_segment(on_tri, cap.a, cap.b);
            on_tri = closest_on_triangle(on_shape, v0, v1, v2);
        }
    }
    const float radius = s.type == SH_SPHERE ? s.radius : s.cr;
    const float pen = radius - len(on_shape - on_tri);
    const float core_side = dot(on_shape - v0, fn);
    if (core_side < 0.0f) {
        const v3 under = on_shape - core_side * fn;
        if (outside(barycentric(under, v0, v1, v2))) return;
        const float exit = radius - core_side;
        add_point(out, under + fn * (exit * 0.5f), exit, tri);
        last_n = fn;
        return;
    }
    if (!(pen > 0.0f)) return;
    const v3 diff = on_shape - on_tri;
    const float dist = len(diff);
    const v3 raw = dist > 1e-6f ? diff / dist : fn;
    const int edge = tri_edge(barycentric(on_tri, v0, v1, v2));
    v3 normal;
    if (edge < 0) {
        normal = dot(raw, fn) > 0.0f ? fn : raw;
    } else {
        const uint4 a = t.adj[mp.meta.y + tri];
        const uint32_t nt = edge == 0 ? a.x : (edge == 1 ? a.y : a.z);
        if (nt == kMeshNone) {
            normal = raw;
        } else {
            const v3 nn = normalize(qrot(mp.rot, ld3(t.n[mp.meta.y + nt])));
            if (dot(fn, nn) > 0.999f) normal = fn;
            else normal = dot(raw, fn) >= dot(raw, nn) ? fn : nn;
        }
    }
    add_point(out, on_tri + normal * (pen * 0.5f), pen, tri);
    last_n = normal;
}

// The reduced manifold of generate_box_mesh_contacts_ / generate_hull_mesh_contacts_
// (mesh_contact.h:565-614): candidates are the valid corners in corner order, then the mesh
// vertices in the order found. Corner c: point corner(c) + n_c * (pen_c / 2); vertex j: point
// key_j - n_j * (depth_j / 2).
struct CornerSet {
    uint64_t valid;
    float pen[64];
    v3 n[64];
};
struct VertexSet {
    int count;
    v3 key[kMeshVerts], n[kMeshVerts];
    float depth[kMeshVerts];
    uint32_t fid[kMeshVerts];
};
struct Recovery {
    bool valid;
    v3 under, n;
    float exit;
    uint32_t fid;
};
template <class CornerAt>
PXD bool reduce_candidates(const CornerSet& cs, int n_corners, const VertexSet& vs, const Recovery& rec,
                           CornerAt&& corner_at, Man& out) {
    int cidx[64];
    int nc = 0;
    for (int c = 0; c < n_corners; ++c)
        if ((cs.valid >> c) & 1ull) cidx[nc++] = c;
    const int total = nc + vs.count;
    if (total == 0) {
        if (!rec.valid) return false;
        add_point(out, rec.under, rec.exit, rec.fid);
        out.n = rec.n;
        return true;
    }
    auto point = [&](int i) -> v3 {
        if (i < nc) {
            const int c = cidx[i];
            return corner_at(c) + cs.n[c] * (cs.pen[c] * 0.5f);
        }
        const int j = i - nc;
        return vs.key[j] - vs.n[j] * (vs.depth[j] * 0.5f);
    };
    auto pen = [&](int i) -> float { return i < nc ? cs.pen[cidx[i]] : vs.depth[i - nc]; };
    auto nrm = [&](int i) -> v3 { return i < nc ? cs.n[cidx[i]] : vs.n[i - nc]; };
    auto fid = [&](int i) -> uint32_t { return i < nc ? uint32_t(cidx[i]) : vs.fid[i - nc]; };
    int pick[4] = {0, 0, 0, 0};
    int picked = 0;
    int deepest = 0;
    for (int i = 1; i < total; ++i)
        if (pen(i) > pen(deepest) + 1e-6f) deepest = i;
    pick[picked++] = deepest;
    const v3 p0 = point(pick[0]);
    for (int round = 0; round < 3 && picked == round + 1; ++round) {
        const v3 p1 = round >= 1 ? point(pick[1]) : p0, p2 = round >= 2 ? point(pick[2]) : p0;
        int arg = total;
        float top = -1.0f;
        for (int i = 0; i < total; ++i) {
            bool taken = false;
            for (int q = 0; q < picked; ++q) taken = taken || pick[q] == i;
            if (taken) continue;
            const v3 p = point(i);
            float sc;
            if (round == 0) {
                sc = len(p - p0);
            } else if (round == 1) {
                sc = len(cross(p1 - p0, p - p0));
            } else {
                sc = len(p - p0);
                const float b = len(p - p1), d = len(p - p2);
                if (b < sc) sc = b;
                if (d < sc) sc = d;
            }
            if (sc > top + 1e-6f) {
                top = sc;
                arg = i;
            }
        }
        if (arg < total && top > 1e-6f) pick[picked++] = arg;
    }
    for (int i = 1; i < picked; ++i) {  // std::sort of at most 4 distinct indices
        const int x = pick[i];
        int j = i;
        while (j > 0 && pick[j - 1] > x) {
            pick[j] = pick[j - 1];
            --j;
        }
        pick[j] = x;
    }
    v3 normal = splat(0.0f);
    for (int i = 0; i < picked; ++i) {
        add_point(out, point(pick[i]), pen(pick[i]), fid(pick[i]));
        normal = normal + nrm(pick[i]) * gmax(pen(pick[i]), 1e-4f);
    }
    const float l = len(normal);
    out.n = l > 1e-6f ? normal / l : nrm(pick[0]);
    return true;
}

PXD void add_vertex(VertexSet& vs, v3 key, v3 n, float depth, uint32_t fid, float inside_eps, bool& overflow) {
    for (int j = 0; j < vs.count; ++j)
        if (len(vs.key[j] - key) < inside_eps) return;
    if (vs.count == kMeshVerts) {
        overflow = true;
        return;
    }
    vs.key[vs.count] = key; vs.n[vs.count] = n; vs.depth[vs.count] = depth; vs.fid[vs.count] = fid;
    ++vs.count;
}
PXD void keep_recovery(Recovery& rec, v3 under, v3 n, float exit, uint32_t tri) {
    if (rec.valid && rec.exit <= exit) return;
    rec.valid = true; rec.under = under; rec.n = n; rec.exit = exit; rec.fid = 0x80000000u | tri;
}

// generate_box_mesh_contacts_ (mesh_contact.h:471-615)
PXD bool box_mesh(const MeshTab& t, const MeshPose& mp, v3 qlo, v3 qhi, const OBB& box, Man& out, bool& overflow) {
    v3 ax[3];
    axes(box, ax);
    const v3 he = box.he;
    v3 corner[8];
    for (uint32_t c = 0; c < 8; ++c)
        corner[c] = box.c + ((c & 1u) ? he.x : -he.x) * ax[0] + ((c & 2u) ? he.y : -he.y) * ax[1] +
                    ((c & 4u) ? he.z : -he.z) * ax[2];
    const float inside_eps = 1e-4f * gmax(1.0f, len(he));
    CornerSet cs;
    cs.valid = 0;
    VertexSet vs;
    vs.count = 0;
    Recovery rec;
    rec.valid = false;
    mesh_candidates(t, mp.meta, qlo, qhi, [&](uint32_t tri) {
        const WTri w = world_tri(t, mp, tri);
        const v3 n = w.n;
        const float center_dist = dot(box.c - w.v[0], n);
        if (center_dist <= 0.0f) {
            const v3 under = box.c - center_dist * n;
            if (outside(barycentric(under, w.v[0], w.v[1], w.v[2]))) return true;
            const float support = fabsf(dot(he.x * ax[0], n)) + fabsf(dot(he.y * ax[1], n)) + fabsf(dot(he.z * ax[2], n));
            keep_recovery(rec, under, n, support - center_dist, tri);
            return true;
        }
        float depth = 0.0f;
        float dist[8];
        for (int c = 0; c < 8; ++c) {
            dist[c] = dot(corner[c] - w.v[0], n);
            depth = gmax(depth, -dist[c]);
        }
        if (depth <= 0.0f) return true;
        for (int c = 0; c < 8; ++c) {
            if (dist[c] >= 0.0f) continue;
            if (outside(barycentric(corner[c] - dist[c] * n, w.v[0], w.v[1], w.v[2]))) continue;
            const float pen = -dist[c];
            if (((cs.valid >> c) & 1ull) && cs.pen[c] <= pen) continue;
            cs.valid |= 1ull << c;
            cs.pen[c] = pen;
            cs.n[c] = n;
        }
        for (uint32_t k = 0; k < 3; ++k) {
            const v3 d = w.v[k] - box.c;
            if (fabsf(dot(d, ax[0])) > he.x + inside_eps || fabsf(dot(d, ax[1])) > he.y + inside_eps ||
                fabsf(dot(d, ax[2])) > he.z + inside_eps)
                continue;
            add_vertex(vs, w.v[k], n, depth, 8u + tri * 3u + k, inside_eps, overflow);
        }
        return true;
    });
    return reduce_candidates(cs, 8, vs, rec, [&](int c) { return corner[c]; }, out);
}

// generate_hull_mesh_contacts_ (mesh_contact.h:645-821)
PXD bool hull_mesh(const MeshTab& t, const MeshPose& mp, v3 qlo, v3 qhi, const HullTab& ht, const PHull& h, uint32_t hull,
                   Man& out, bool& overflow) {
    const v3 hc = h.pos + qrot(h.rot, ld3(ht.cen[hull]) * h.scale);
    const float inside_eps = 1e-4f * gmax(1.0f, len(ld3(ht.lb[2 * hull + 1]) - ld3(ht.lb[2 * hull])) * h.scale);
    CornerSet cs;
    cs.valid = 0;
    VertexSet vs;
    vs.count = 0;
    Recovery rec;
    rec.valid = false;
    mesh_candidates(t, mp.meta, qlo, qhi, [&](uint32_t tri) {
        const WTri w = world_tri(t, mp, tri);
        const v3 n = w.n;
        const float center_dist = dot(hc - w.v[0], n);
        if (center_dist <= 0.0f) {
            const v3 under = hc - center_dist * n;
            if (outside(barycentric(under, w.v[0], w.v[1], w.v[2]))) return true;
            float support = 0.0f;
            for (uint32_t c = 0; c < h.nv; ++c) support = gmax(support, dot(hvtx(ht, h, c) - hc, n));
            keep_recovery(rec, under, n, support - center_dist, tri);
            return true;
        }
        float depth = 0.0f;
        for (uint32_t c = 0; c < h.nv; ++c) depth = gmax(depth, -dot(hvtx(ht, h, c) - w.v[0], n));
        if (depth <= 0.0f) return true;
        for (uint32_t c = 0; c < h.nv; ++c) {
            const v3 p = hvtx(ht, h, c);
            const float dist = dot(p - w.v[0], n);
            if (dist >= 0.0f) continue;
            if (outside(barycentric(p - dist * n, w.v[0], w.v[1], w.v[2]))) continue;
            const float pen = -dist;
            if (((cs.valid >> c) & 1ull) && cs.pen[c] <= pen) continue;
            cs.valid |= 1ull << c;
            cs.pen[c] = pen;
            cs.n[c] = n;
        }
        for (uint32_t k = 0; k < 3; ++k) {
            bool inside = true;
            for (uint32_t f = 0; f < h.nf && inside; ++f) {
                const v3 fnw = hfn(ht, h, f);
                if (dot(fnw, w.v[k]) - dot(fnw, hface_v0(ht, h, f)) > inside_eps) inside = false;
            }
            if (inside) add_vertex(vs, w.v[k], n, depth, h.nv + tri * 3u + k, inside_eps, overflow);
        }
        return true;
    });
    return reduce_candidates(cs, int(h.nv), vs, rec, [&](int c) { return hvtx(ht, h, uint32_t(c)); }, out);
}

// generate_mesh_contacts (mesh_contact.h:838-884): normal from the mesh toward the shape.
// mesh_contact_candidates (330-340) grows a box's bounds by 1.5 x its half diagonal and a hull's
// by 1.5 x half its local-bounds diagonal x scale before mapping them into the mesh frame.
PXD bool mesh_contacts(const MeshTab& t, const HullTab& ht, const ShapeGeom& s, v3 pos, quat rot, const ShapeGeom& ms,
                       v3 mpos, quat mrot, Man& out, bool& overflow) {
    out.count = 0;
    out.n = mk(0.0f, 0.0f, 1.0f);
    MeshPose mp;
    mp.meta = t.meta[ms.hull];
    mp.mt = mesh_transform(ms, mpos, mrot);
    mp.rot = qmul(mrot, ms.lr);
    v3 lo, hi;
    if (s.type == SH_HULL) hull_bounds(ht, s, pos, rot, lo, hi);
    else world_bounds(s, pos, rot, lo, hi);
    if (s.type == SH_BOX) {
        const float g = 1.5f * len(s.he);
        lo = lo - splat(g);
        hi = hi + splat(g);
    } else if (s.type == SH_HULL) {
        const float g = 0.75f * len(ld3(ht.lb[2 * s.hull + 1]) - ld3(ht.lb[2 * s.hull])) * s.radius;
        lo = lo - splat(g);
        hi = hi + splat(g);
    }
    v3 qlo, qhi;
    box_through(m4_inverse(mp.mt), lo, hi, qlo, qhi);
    if (s.type == SH_BOX) return box_mesh(t, mp, qlo, qhi, world_obb(s, pos, rot), out, overflow);
    if (s.type == SH_HULL) return hull_mesh(t, mp, qlo, qhi, ht, posed_hull(ht, s.hull, pos, rot, s), s.hull, out, overflow);
    v3 last_n = mk(0.0f, 0.0f, 1.0f);
    mesh_candidates(t, mp.meta, qlo, qhi, [&](uint32_t tri) {
        if (out.count >= 4) return false;
        round_tri(t, mp, tri, s, pos, rot, out, last_n);
        return true;
    });
    if (out.count == 0) return false;
    out.n = last_n;
    return true;
}

}  // namespace pxd
