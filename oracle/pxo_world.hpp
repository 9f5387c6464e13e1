// ORACLE -- test infrastructure only. Nothing in the product may include, link or run this.
//
// A CPU restatement of physxcoopa's rigid-body contact step (`PhysicsWorld::step_fixed`,
// physxcoopa/world.h:1278-1317) for box / sphere / capsule shapes, templated on the scalar type:
// T = float is the parity reference for the GPU, T = double is the error reference. Each function
// cites the reference file:line it restates. Two additions that the reference does not have:
//   * a solve order: the valid manifolds can be permuted (the GPU's colour order) before the
//     unchanged sequential solve (dynamics/solver.h:914-958), so a coloured GPU solve can be
//     checked against it bit for bit;
//   * an optional table of initial impulses for warm-start misses (the learned warm start's
//     hook at dynamics/solver.h:166-170), and the velocity residual.
// The broadphase is the reference's dynamic AABB tree (broadphase/aabb_tree.h) with per-proxy
// fat AABBs; a brute-force pair finder with the same pair rule checks it.
#pragma once
#include "pxo_math.hpp"
#include "../include/pxg.h"

#include <vector>
#include <unordered_map>
#include <algorithm>
#include <cstring>

namespace pxo {

template <class T> struct AABB {
    V3<T> mn{std::numeric_limits<T>::max()};
    V3<T> mx{std::numeric_limits<T>::lowest()};
    // geometry/aabb.h:35-39 (inclusive)
    bool overlaps(const AABB& o) const {
        return mn.x <= o.mx.x && mx.x >= o.mn.x && mn.y <= o.mx.y && mx.y >= o.mn.y && mn.z <= o.mx.z && mx.z >= o.mn.z;
    }
    bool contains(const V3<T>& p) const {
        return p.x >= mn.x && p.x <= mx.x && p.y >= mn.y && p.y <= mx.y && p.z >= mn.z && p.z <= mx.z;
    }
    bool contains(const AABB& o) const { return contains(o.mn) && contains(o.mx); }  // aabb.h:52-54
    AABB expand(T m) const { AABB r; r.mn = mn - V3<T>(m); r.mx = mx + V3<T>(m); return r; }  // aabb.h:57-62
    static AABB merge(const AABB& a, const AABB& b) { AABB r; r.mn = vmin(a.mn, b.mn); r.mx = vmax(a.mx, b.mx); return r; }
    T area() const {  // aabb.h:65-70
        V3<T> e = mx - mn;
        if (e.x < T(0) || e.y < T(0) || e.z < T(0)) return T(0);
        return T(2) * (e.x * e.y + e.y * e.z + e.z * e.x);
    }
};

template <class T> struct Body {
    int type = PXG_DYNAMIC;
    bool alive = true, awake = true, use_gravity = true;
    uint32_t constraints = 0, generation = 0;
    V3<T> pos, v, w, force, torque, iil{T(1)};
    Q<T> q;
    T mass = 1, inv_mass = 1, ldrag = 0, adrag = T(0.05f), sleep_timer = 0;
    M3<T> iiw{T(1)};
    V3<T> prev_pos;
    Q<T> prev_q;
};

template <class T> struct Shape {
    bool enabled = false, trigger = false;
    int type = PXG_SPHERE, caxis = 2;
    uint32_t owner = 0, layer = 0;
    V3<T> lc;
    Q<T> lr;
    T radius = T(0.5f), cr = T(0.5f), chh = T(0.5f);
    V3<T> he{T(0.5f)};
    T friction = T(0.6f), restitution = 0;
    int fcomb = 0, rcomb = 0;
    const void* hull = nullptr;  // PXG_CONVEX_HULL: the world's Hull<T> (shape.h:97)
    T hull_scale = T(1);
    const void* mesh = nullptr;  // PXG_TRIANGLE_MESH: the world's Mesh (shape.h:85)
    T mesh_scale = T(1);
};

template <class T> struct Contact {
    V3<T> pos;
    T pen = 0, sep = 0, ni = 0, ti[2] = {0, 0}, rbias = 0;
    uint32_t fid = 0;
};

template <class T> struct Manifold {
    uint32_t a = 0, b = 0, sa = 0, sb = 0;
    V3<T> n{T(0), T(0), T(1)};
    Contact<T> p[4];
    int count = 0;
    T friction = 0, restitution = 0;
    bool trigger = false, valid = false;
    void add_point(const V3<T>& pos, T pen, uint32_t fid) {  // collision/manifold.h:78-84
        if (count >= 4) return;
        Contact<T>& c = p[count++];
        c = Contact<T>{};
        c.pos = pos; c.pen = pen; c.fid = fid;
    }
    void add_spec(const V3<T>& pos, T sep, uint32_t fid) {  // collision/manifold.h:89-95
        if (count >= 4) return;
        Contact<T>& c = p[count++];
        c = Contact<T>{};
        c.pos = pos; c.sep = sep; c.fid = fid;
    }
};

struct Cached { double n = 0, t1 = 0, t2 = 0; };

// ---------------------------------------------------------------------------------------------
// geometry (geometry/obb.h, geometry/capsule.h, collision/shape.h)
// ---------------------------------------------------------------------------------------------
template <class T> struct OBB { V3<T> c; V3<T> he; Q<T> q; };
template <class T> struct Sph { V3<T> c; T r; };
template <class T> struct Cap { V3<T> a, b; T r; };

template <class T> inline void obb_axes(const OBB<T>& o, V3<T> ax[3]) {  // obb.h:31-37
    M3<T> r = mat3_cast(o.q);
    ax[0] = r.c[0]; ax[1] = r.c[1]; ax[2] = r.c[2];
}
template <class T> inline AABB<T> obb_bounds(const OBB<T>& o) {  // obb.h:39-49 + aabb.h:83-95
    M3<T> m = mat3_cast(o.q);
    AABB<T> out;
    for (int i = 0; i < 8; ++i) {
        V3<T> corner((i & 1) ? o.he.x : -o.he.x, (i & 2) ? o.he.y : -o.he.y, (i & 4) ? o.he.z : -o.he.z);
        V3<T> wpt = affine(m, o.c, corner);
        out.mn = vmin(out.mn, wpt);
        out.mx = vmax(out.mx, wpt);
    }
    return out;
}
template <class T> inline V3<T> closest_point_on_obb(const V3<T>& p, const OBB<T>& b) {  // obb.h:65-75
    V3<T> ax[3];
    obb_axes(b, ax);
    V3<T> d = p - b.c;
    T px = sclamp(dot(d, ax[0]), -b.he.x, b.he.x);
    T py = sclamp(dot(d, ax[1]), -b.he.y, b.he.y);
    T pz = sclamp(dot(d, ax[2]), -b.he.z, b.he.z);
    return b.c + ax[0] * px + ax[1] * py + ax[2] * pz;
}
template <class T> inline V3<T> closest_point_on_segment(const V3<T>& p, const V3<T>& a, const V3<T>& b) {  // capsule.h:30-40
    V3<T> ab = b - a;
    T len2 = dot(ab, ab);
    if (len2 < k_epsilon<T>()) return a;
    T t = dot(p - a, ab) / len2;
    t = sclamp(t, T(0), T(1));
    return a + ab * t;
}
template <class T> inline void closest_points_seg_seg(const V3<T>& p1, const V3<T>& q1, const V3<T>& p2, const V3<T>& q2,
                                                      V3<T>& c1, V3<T>& c2) {  // capsule.h:73-134
    V3<T> d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    T s, t;
    const T eps = k_epsilon<T>();
    if (a < eps && e < eps) { c1 = p1; c2 = p2; return; }
    if (a < eps) {
        s = 0;
        t = sclamp(f / e, T(0), T(1));
    } else {
        T c = dot(d1, r);
        if (e < eps) {
            t = 0;
            s = sclamp(-c / a, T(0), T(1));
        } else {
            T b = dot(d1, d2);
            T denom = a * e - b * b;
            if (denom > eps) s = sclamp((b * f - c * e) / denom, T(0), T(1));
            else s = 0;
            t = (b * s + f) / e;
            if (t < T(0)) { t = 0; s = sclamp(-c / a, T(0), T(1)); }
            else if (t > T(1)) { t = 1; s = sclamp((b - c) / a, T(0), T(1)); }
        }
    }
    c1 = p1 + d1 * s;
    c2 = p2 + d2 * t;
}

template <class T> inline Sph<T> world_sphere(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot) {  // shape.h:160-165
    return Sph<T>{pos + qrot(rot, s.lc), s.radius};
}
template <class T> inline OBB<T> world_obb(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot) {  // shape.h:168-174
    return OBB<T>{pos + qrot(rot, s.lc), s.he, qmul(rot, s.lr)};
}
template <class T> inline Cap<T> world_capsule(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot) {  // shape.h:177-187
    V3<T> axl(0);
    axl[s.caxis] = T(1);
    V3<T> axw = qrot(qmul(rot, s.lr), axl);
    V3<T> c = pos + qrot(rot, s.lc);
    return Cap<T>{c - axw * s.chh, c + axw * s.chh, s.cr};
}
template <class T> AABB<T> hull_world_bounds(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot);  // pxo_hull.hpp
template <class T> AABB<T> mesh_world_bounds(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot);  // pxo_mesh.hpp
template <class T> bool mesh_pair_contacts(const Shape<T>& sa, const V3<T>& pa, const Q<T>& ra, const Shape<T>& sb,
                                           const V3<T>& pb, const Q<T>& rb, Manifold<T>& out, bool& hit);
template <class T> T hull_min_extent(const Shape<T>& s);
template <class T> bool hull_pair_contacts(const Shape<T>& sa, const V3<T>& pa, const Q<T>& ra, const Shape<T>& sb,
                                           const V3<T>& pb, const Q<T>& rb, Manifold<T>& out, bool& hit);
template <class T> inline AABB<T> world_bounds(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot) {  // shape.h:198-228
    AABB<T> out;
    if (s.type == PXG_CONVEX_HULL) return hull_world_bounds(s, pos, rot);
    if (s.type == PXG_TRIANGLE_MESH) return mesh_world_bounds(s, pos, rot);
    if (s.type == PXG_SPHERE) {
        Sph<T> sp = world_sphere(s, pos, rot);
        out.mn = sp.c - V3<T>(sp.r);
        out.mx = sp.c + V3<T>(sp.r);
    } else if (s.type == PXG_BOX) {
        out = obb_bounds(world_obb(s, pos, rot));
    } else {
        Cap<T> c = world_capsule(s, pos, rot);
        out.mn = vmin(c.a, c.b) - V3<T>(c.r);
        out.mx = vmax(c.a, c.b) + V3<T>(c.r);
    }
    return out;
}

// ---------------------------------------------------------------------------------------------
// narrowphase (collision/narrowphase.h, sat.h, clip.h, segment.h)
// ---------------------------------------------------------------------------------------------
template <class T> bool sphere_vs_sphere(const Sph<T>& a, const Sph<T>& b, V3<T>& n, T& pen, V3<T>& pt) {  // narrowphase.h:40-52
    V3<T> d = b.c - a.c;
    T d2 = dot(d, d);
    T rs = a.r + b.r;
    if (d2 > rs * rs) return false;
    T dist = std::sqrt(d2);
    n = dist > k_epsilon<T>() ? d / dist : V3<T>(0, 0, 1);
    pen = rs - dist;
    pt = a.c + n * a.r;
    return true;
}
template <class T> bool sphere_vs_box(const Sph<T>& s, const OBB<T>& b, V3<T>& n, T& pen, V3<T>& pt) {  // narrowphase.h:64-101
    V3<T> ax[3];
    obb_axes(b, ax);
    V3<T> d = s.c - b.c;
    V3<T> local(dot(d, ax[0]), dot(d, ax[1]), dot(d, ax[2]));
    bool inside = std::abs(local.x) <= b.he.x && std::abs(local.y) <= b.he.y && std::abs(local.z) <= b.he.z;
    V3<T> nl, cl;
    if (!inside) {
        cl = V3<T>(gclamp(local.x, -b.he.x, b.he.x), gclamp(local.y, -b.he.y, b.he.y), gclamp(local.z, -b.he.z, b.he.z));
        V3<T> diff = local - cl;
        T dist = length(diff);
        if (dist > s.r) return false;
        pen = s.r - dist;
        nl = dist > k_epsilon<T>() ? diff / dist : V3<T>(0, 0, 1);
    } else {
        V3<T> fd(b.he.x - gabs(local.x), b.he.y - gabs(local.y), b.he.z - gabs(local.z));
        int axis = 0;
        if (fd.y < fd.x) axis = 1;
        if (fd.z < (axis == 0 ? fd.x : fd.y)) axis = 2;
        nl = V3<T>(0);
        nl[axis] = local[axis] >= T(0) ? T(1) : T(-1);
        pen = fd[axis] + s.r;
        cl = local;
        cl[axis] = nl[axis] * b.he[axis];
    }
    n = ax[0] * nl.x + ax[1] * nl.y + ax[2] * nl.z;
    pt = b.c + ax[0] * cl.x + ax[1] * cl.y + ax[2] * cl.z;
    return true;
}
template <class T> bool sphere_vs_sphere_sep(const Sph<T>& a, const Sph<T>& b, V3<T>& n, T& sep, V3<T>& pt) {  // narrowphase.h:112-122
    V3<T> d = b.c - a.c;
    T dist = length(d);
    T rs = a.r + b.r;
    if (dist <= rs || dist < k_epsilon<T>()) return false;
    n = d / dist;
    sep = dist - rs;
    pt = a.c + n * a.r;
    return true;
}
template <class T> bool sphere_vs_box_sep(const Sph<T>& s, const OBB<T>& b, V3<T>& n, T& sep, V3<T>& pt) {  // narrowphase.h:133-154
    V3<T> ax[3];
    obb_axes(b, ax);
    V3<T> d = s.c - b.c;
    V3<T> local(dot(d, ax[0]), dot(d, ax[1]), dot(d, ax[2]));
    bool inside = std::abs(local.x) <= b.he.x && std::abs(local.y) <= b.he.y && std::abs(local.z) <= b.he.z;
    if (inside) return false;
    V3<T> cl(gclamp(local.x, -b.he.x, b.he.x), gclamp(local.y, -b.he.y, b.he.y), gclamp(local.z, -b.he.z, b.he.z));
    V3<T> diff = local - cl;
    T dist = length(diff);
    if (dist <= s.r) return false;
    sep = dist - s.r;
    V3<T> nl = diff / dist;
    n = ax[0] * nl.x + ax[1] * nl.y + ax[2] * nl.z;
    pt = b.c + ax[0] * cl.x + ax[1] * cl.y + ax[2] * cl.z;
    return true;
}

template <class T> struct SatResult { T pen = 0, sep = 0; int axis = -1; V3<T> n{T(0), T(0), T(1)}; bool separated = false; };

template <class T> bool sat_obb_obb(const OBB<T>& a, const OBB<T>& b, SatResult<T>& res, bool allow_sep) {  // sat.h:75-160
    V3<T> aa[3], ba[3];
    obb_axes(a, aa);
    obb_axes(b, ba);
    V3<T> d = b.c - a.c;
    T best_score = std::numeric_limits<T>::max(), best_overlap = 0;
    int best_axis = -1;
    V3<T> best_n(0, 0, 1);
    bool separated = false;
    T best_sep_overlap = -std::numeric_limits<T>::max();
    int best_sep_axis = -1;
    V3<T> best_sep_n(0, 0, 1);
    auto test_axis = [&](const V3<T>& raw, int id, T bias) {
        if (separated && !allow_sep) return;
        T len2 = dot(raw, raw);
        if (len2 < T(1e-8f)) return;
        V3<T> axis = raw / std::sqrt(len2);
        T ra = 0, rb = 0;
        for (int k = 0; k < 3; ++k) {
            ra += a.he[k] * std::abs(dot(axis, aa[k]));
            rb += b.he[k] * std::abs(dot(axis, ba[k]));
        }
        T dist = std::abs(dot(d, axis));
        T overlap = ra + rb - dist;
        V3<T> n = dot(d, axis) < T(0) ? -axis : axis;
        if (overlap < T(0)) {
            separated = true;
            if (allow_sep && overlap > best_sep_overlap) { best_sep_overlap = overlap; best_sep_axis = id; best_sep_n = n; }
            return;
        }
        T score = overlap + bias;
        if (score < best_score) { best_score = score; best_overlap = overlap; best_axis = id; best_n = n; }
    };
    const T edge_bias = T(0.01f);
    test_axis(aa[0], 0, T(0)); test_axis(aa[1], 1, T(0)); test_axis(aa[2], 2, T(0));
    test_axis(ba[0], 3, T(0)); test_axis(ba[1], 4, T(0)); test_axis(ba[2], 5, T(0));
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            if (separated && !allow_sep) break;
            test_axis(cross(aa[i], ba[j]), 6 + i * 3 + j, edge_bias);
        }
        if (separated && !allow_sep) break;
    }
    if (separated) {
        if (!allow_sep || best_sep_axis < 0) return false;
        res.pen = 0; res.sep = -best_sep_overlap; res.axis = best_sep_axis; res.n = best_sep_n; res.separated = true;
        return true;
    }
    if (best_axis < 0) return false;
    res.pen = best_overlap; res.sep = 0; res.axis = best_axis; res.n = best_n; res.separated = false;
    return true;
}

template <class T> struct Plane { V3<T> n; T d; };
constexpr int k_max_clip = 9;
template <class T> void clip_polygon(const V3<T>* in, int nin, const Plane<T>& pl, V3<T>* out, int& nout) {  // clip.h:40-64
    nout = 0;
    if (nin == 0) return;
    for (int i = 0; i < nin; ++i) {
        const V3<T>& cur = in[i];
        const V3<T>& prev = in[(i - 1 + nin) % nin];
        T dc = dot(pl.n, cur) - pl.d;
        T dp = dot(pl.n, prev) - pl.d;
        bool ci = dc <= T(0), pi = dp <= T(0);
        if (ci != pi) {
            T t = dp / (dp - dc);
            if (nout < k_max_clip) out[nout++] = prev + t * (cur - prev);
        }
        if (ci) { if (nout < k_max_clip) out[nout++] = cur; }
    }
}
template <class T> void obb_face(const OBB<T>& b, int axis, T sign, V3<T> corners[4], V3<T>& normal) {  // clip.h:75-92
    V3<T> ax[3];
    obb_axes(b, ax);
    const V3<T>& m = ax[axis];
    const V3<T>& u = ax[(axis + 1) % 3];
    const V3<T>& v = ax[(axis + 2) % 3];
    T hu = b.he[(axis + 1) % 3], hv = b.he[(axis + 2) % 3];
    V3<T> fc = b.c + m * (sign * b.he[axis]);
    corners[0] = fc - u * hu - v * hv;
    corners[1] = fc + u * hu - v * hv;
    corners[2] = fc + u * hu + v * hv;
    corners[3] = fc - u * hu + v * hv;
    normal = m * sign;
}
template <class T> void obb_side_planes(const OBB<T>& b, int axis, Plane<T> pl[4]) {  // clip.h:102-118
    V3<T> ax[3];
    obb_axes(b, ax);
    const V3<T>& u = ax[(axis + 1) % 3];
    const V3<T>& v = ax[(axis + 2) % 3];
    T hu = b.he[(axis + 1) % 3], hv = b.he[(axis + 2) % 3];
    pl[0] = Plane<T>{u, dot(u, b.c) + hu};
    pl[1] = Plane<T>{-u, dot(-u, b.c) + hu};
    pl[2] = Plane<T>{v, dot(v, b.c) + hv};
    pl[3] = Plane<T>{-v, dot(-v, b.c) + hv};
}

template <class T> bool box_box(const OBB<T>& a, const OBB<T>& b, Manifold<T>& out, bool allow_spec) {  // sat.h:190-331
    SatResult<T> r;
    if (!sat_obb_obb(a, b, r, allow_spec)) return false;
    out = Manifold<T>{};
    out.n = r.n;
    if (r.separated) {
        V3<T> ca = closest_point_on_obb(b.c, a);
        V3<T> cb = closest_point_on_obb(a.c, b);
        V3<T> pt = T(0.5f) * (ca + cb);
        out.add_spec(pt, r.sep, 0x2000u | static_cast<uint32_t>(r.axis));
        return true;
    }
    if (r.axis <= 5) {
        bool ref_is_a = r.axis <= 2;
        const OBB<T>& ref = ref_is_a ? a : b;
        const OBB<T>& inc = ref_is_a ? b : a;
        int ref_axis = ref_is_a ? r.axis : r.axis - 3;
        V3<T> rax[3];
        obb_axes(ref, rax);
        V3<T> desired = ref_is_a ? r.n : -r.n;
        T ref_sign = dot(rax[ref_axis], desired) >= T(0) ? T(1) : T(-1);
        V3<T> rc[4], rn;
        obb_face(ref, ref_axis, ref_sign, rc, rn);
        V3<T> iax[3];
        obb_axes(inc, iax);
        int inc_axis = 0;
        T inc_sign = 1;
        T most = std::numeric_limits<T>::max();
        for (int k = 0; k < 3; ++k) {
            for (int si = 0; si < 2; ++si) {
                T s = si == 0 ? T(1) : T(-1);
                T dp = dot(iax[k] * s, rn);
                if (dp < most) { most = dp; inc_axis = k; inc_sign = s; }
            }
        }
        V3<T> ic[4], inn;
        obb_face(inc, inc_axis, inc_sign, ic, inn);
        Plane<T> pl[4];
        obb_side_planes(ref, ref_axis, pl);
        V3<T> pa[k_max_clip], pb[k_max_clip];
        int na = 4;
        for (int i = 0; i < 4; ++i) pa[i] = ic[i];
        for (int p = 0; p < 4 && na > 0; ++p) {
            int nb = 0;
            clip_polygon(pa, na, pl[p], pb, nb);
            for (int i = 0; i < nb; ++i) pa[i] = pb[i];
            na = nb;
        }
        if (na == 0) return false;
        T ref_d = dot(rn, rc[0]);
        struct Cand { V3<T> pos; T pen; };
        Cand cand[k_max_clip];
        int kept = 0;
        for (int i = 0; i < na; ++i) {
            T sd = dot(rn, pa[i]) - ref_d;
            T pen = -sd;
            if (pen > -k_epsilon<T>()) { cand[kept].pos = pa[i]; cand[kept].pen = gmax(pen, T(0)); ++kept; }
        }
        if (kept == 0) return false;
        if (kept > 4) {
            // libstdc++ std::sort on n <= 16 elements is an insertion sort (stable): restate it as a
            // stable insertion sort by penetration, descending.
            for (int i = 1; i < kept; ++i) {
                Cand c = cand[i];
                int j = i;
                while (j > 0 && c.pen > cand[j - 1].pen) { cand[j] = cand[j - 1]; --j; }
                cand[j] = c;
            }
            kept = 4;
        }
        int rfid = ref_axis * 2 + (ref_sign < T(0) ? 1 : 0);
        int ifid = inc_axis * 2 + (inc_sign < T(0) ? 1 : 0);
        for (int i = 0; i < kept; ++i) {
            uint32_t fid = (uint32_t(rfid) << 8) | (uint32_t(ifid) << 4) | uint32_t(i & 0xF);
            out.add_point(cand[i].pos, cand[i].pen, fid);
        }
        return out.count > 0;
    }
    int idx = r.axis - 6, ei = idx / 3, ej = idx % 3;
    V3<T> aa[3], ba[3];
    obb_axes(a, aa);
    obb_axes(b, ba);
    V3<T> dab = b.c - a.c;
    int aj = (ei + 1) % 3, ak = (ei + 2) % 3;
    T asj = dot(dab, aa[aj]) >= T(0) ? T(1) : T(-1);
    T ask = dot(dab, aa[ak]) >= T(0) ? T(1) : T(-1);
    V3<T> eac = a.c + aa[aj] * (asj * a.he[aj]) + aa[ak] * (ask * a.he[ak]);
    V3<T> ea0 = eac - aa[ei] * a.he[ei], ea1 = eac + aa[ei] * a.he[ei];
    V3<T> dba = -dab;
    int bj = (ej + 1) % 3, bk = (ej + 2) % 3;
    T bsj = dot(dba, ba[bj]) >= T(0) ? T(1) : T(-1);
    T bsk = dot(dba, ba[bk]) >= T(0) ? T(1) : T(-1);
    V3<T> ebc = b.c + ba[bj] * (bsj * b.he[bj]) + ba[bk] * (bsk * b.he[bk]);
    V3<T> eb0 = ebc - ba[ej] * b.he[ej], eb1 = ebc + ba[ej] * b.he[ej];
    V3<T> c1, c2;
    closest_points_seg_seg(ea0, ea1, eb0, eb1, c1, c2);
    V3<T> pt = T(0.5f) * (c1 + c2);
    out.add_point(pt, r.pen, 0x1000u | (uint32_t(ei) << 4) | uint32_t(ej));
    return true;
}

template <class T> void closest_seg_obb(const V3<T>& a, const V3<T>& b, const OBB<T>& box, V3<T>& on_seg, V3<T>& on_box) {  // segment.h:39-46
    on_seg = T(0.5f) * (a + b);
    for (int i = 0; i < 8; ++i) {
        on_box = closest_point_on_obb(on_seg, box);
        on_seg = closest_point_on_segment(on_box, a, b);
    }
}
template <class T> bool capsule_vs_sphere(const Cap<T>& c, const Sph<T>& s, V3<T>& n, T& pen, V3<T>& pt) {  // segment.h:51-63
    V3<T> cl = closest_point_on_segment(s.c, c.a, c.b);
    V3<T> d = s.c - cl;
    T d2 = dot(d, d);
    T rs = c.r + s.r;
    if (d2 > rs * rs) return false;
    T dist = std::sqrt(d2);
    n = dist > k_epsilon<T>() ? d / dist : V3<T>(0, 0, 1);
    pen = rs - dist;
    pt = cl + n * c.r;
    return true;
}
template <class T> bool capsule_vs_capsule(const Cap<T>& a, const Cap<T>& b, V3<T>& n, T& pen, V3<T>& pt) {  // segment.h:68-81
    V3<T> c1, c2;
    closest_points_seg_seg(a.a, a.b, b.a, b.b, c1, c2);
    V3<T> d = c2 - c1;
    T d2 = dot(d, d);
    T rs = a.r + b.r;
    if (d2 > rs * rs) return false;
    T dist = std::sqrt(d2);
    n = dist > k_epsilon<T>() ? d / dist : V3<T>(0, 0, 1);
    pen = rs - dist;
    pt = c1 + n * a.r;
    return true;
}
template <class T> bool capsule_vs_box(const Cap<T>& cap, const OBB<T>& box, Manifold<T>& out) {  // segment.h:91-136
    V3<T> os, ob;
    closest_seg_obb(cap.a, cap.b, box, os, ob);
    V3<T> delta = os - ob;
    T dist = length(delta);
    if (dist > cap.r) return false;
    V3<T> pn = dist > k_epsilon<T>() ? delta / dist : V3<T>(0, 0, 1);
    T ppen = cap.r - dist;
    out = Manifold<T>{};
    out.n = pn;
    V3<T> ba_ = closest_point_on_obb(cap.a, box);
    V3<T> bb_ = closest_point_on_obb(cap.b, box);
    T da = length(cap.a - ba_);
    T db = length(cap.b - bb_);
    const T flat = T(0.02f);
    bool at = da <= cap.r + flat, bt = db <= cap.r + flat;
    if (at && bt) {
        V3<T> na = cap.a - ba_, nb = cap.b - bb_;
        T nal = length(na), nbl = length(nb);
        V3<T> n_a = nal > k_epsilon<T>() ? na / nal : pn;
        V3<T> n_b = nbl > k_epsilon<T>() ? nb / nbl : pn;
        if (dot(n_a, n_b) > T(0.9f)) {
            out.n = normalize(n_a + n_b);
            out.add_point(ba_, gmax(cap.r - da, T(0)), 0);
            out.add_point(bb_, gmax(cap.r - db, T(0)), 1);
            return true;
        }
    }
    out.add_point(ob, ppen, 0);
    return true;
}

template <class T> inline T combine(T a, T b, int ma, int mb) {  // physics_material.h:70-79
    int m = std::max(ma, mb);
    switch (m) {
        case PXG_MULTIPLY: return a * b;
        case PXG_MAXIMUM: return gmax(a, b);  // std::max(a,b) = (a < b) ? b : a
        case PXG_MINIMUM: return gmin(a, b);  // std::min(a,b) = (b < a) ? b : a
        default: return T(0.5f) * (a + b);
    }
}

// collision/narrowphase.h:175-335 (box/sphere/capsule branches)
template <class T> bool generate_contacts(uint32_t ida, const Shape<T>& sa, const V3<T>& pa, const Q<T>& ra,
                                          uint32_t idb, const Shape<T>& sb, const V3<T>& pb, const Q<T>& rb,
                                          Manifold<T>& out, bool spec) {
    out = Manifold<T>{};
    out.a = ida; out.b = idb;
    bool hit = false;
    V3<T> n, p;
    T pen;
    int ta = sa.type, tb = sb.type;
    if (ta == PXG_SPHERE && tb == PXG_SPHERE) {
        Sph<T> A = world_sphere(sa, pa, ra), B = world_sphere(sb, pb, rb);
        if (sphere_vs_sphere(A, B, n, pen, p)) { out.n = n; out.add_point(p, pen, 0); hit = true; }
        else if (spec && sphere_vs_sphere_sep(A, B, n, pen, p)) { out.n = n; out.add_spec(p, pen, 0); hit = true; }
    } else if (ta == PXG_SPHERE && tb == PXG_BOX) {
        Sph<T> A = world_sphere(sa, pa, ra);
        OBB<T> B = world_obb(sb, pb, rb);
        if (sphere_vs_box(A, B, n, pen, p)) { out.n = -n; out.add_point(p, pen, 0); hit = true; }
        else if (spec && sphere_vs_box_sep(A, B, n, pen, p)) { out.n = -n; out.add_spec(p, pen, 0); hit = true; }
    } else if (ta == PXG_BOX && tb == PXG_SPHERE) {
        OBB<T> A = world_obb(sa, pa, ra);
        Sph<T> B = world_sphere(sb, pb, rb);
        if (sphere_vs_box(B, A, n, pen, p)) { out.n = n; out.add_point(p, pen, 0); hit = true; }
        else if (spec && sphere_vs_box_sep(B, A, n, pen, p)) { out.n = n; out.add_spec(p, pen, 0); hit = true; }
    } else if (ta == PXG_BOX && tb == PXG_BOX) {
        hit = box_box(world_obb(sa, pa, ra), world_obb(sb, pb, rb), out, spec);
    } else if (ta == PXG_CAPSULE && tb == PXG_CAPSULE) {
        if (capsule_vs_capsule(world_capsule(sa, pa, ra), world_capsule(sb, pb, rb), n, pen, p)) { out.n = n; out.add_point(p, pen, 0); hit = true; }
    } else if (ta == PXG_CAPSULE && tb == PXG_SPHERE) {
        if (capsule_vs_sphere(world_capsule(sa, pa, ra), world_sphere(sb, pb, rb), n, pen, p)) { out.n = n; out.add_point(p, pen, 0); hit = true; }
    } else if (ta == PXG_SPHERE && tb == PXG_CAPSULE) {
        if (capsule_vs_sphere(world_capsule(sb, pb, rb), world_sphere(sa, pa, ra), n, pen, p)) { out.n = -n; out.add_point(p, pen, 0); hit = true; }
    } else if (ta == PXG_CAPSULE && tb == PXG_BOX) {
        hit = capsule_vs_box(world_capsule(sa, pa, ra), world_obb(sb, pb, rb), out);
        if (hit) out.n = -out.n;
    } else if (ta == PXG_BOX && tb == PXG_CAPSULE) {
        hit = capsule_vs_box(world_capsule(sb, pb, rb), world_obb(sa, pa, ra), out);
    } else if (ta == PXG_TRIANGLE_MESH || tb == PXG_TRIANGLE_MESH) {
        mesh_pair_contacts(sa, pa, ra, sb, pb, rb, out, hit);
    } else if (!hull_pair_contacts(sa, pa, ra, sb, pb, rb, out, hit)) {
        return false;
    }
    if (!hit) return false;
    out.a = ida; out.b = idb;
    out.friction = combine(sa.friction, sb.friction, sa.fcomb, sb.fcomb);
    out.restitution = combine(sa.restitution, sb.restitution, sa.rcomb, sb.rcomb);
    out.trigger = sa.trigger || sb.trigger;
    out.valid = true;
    return true;
}

// ---------------------------------------------------------------------------------------------
// broadphase: dynamic AABB tree (broadphase/aabb_tree.h:37-305), no rotations
// ---------------------------------------------------------------------------------------------
template <class T> class Tree {
public:
    struct Node { AABB<T> box; int32_t data = -1, parent = -1, c1 = -1, c2 = -1, height = -1; };
    std::vector<Node> nodes;
    int32_t root = -1, free_list = -1;
    T margin = T(0.1f);

    int32_t create(const AABB<T>& tight, int32_t data) {  // aabb_tree.h:45-52
        int32_t p = alloc();
        nodes[p].box = tight.expand(margin);
        nodes[p].data = data;
        nodes[p].height = 0;
        insert(p);
        return p;
    }
    void destroy(int32_t p) { remove(p); free_node(p); }
    bool move(int32_t p, const AABB<T>& tight, const V3<T>& disp) {  // aabb_tree.h:76-85
        AABB<T> sw = tight;
        sw.mn = vmin(sw.mn, tight.mn + disp);
        sw.mx = vmax(sw.mx, tight.mx + disp);
        if (nodes[p].box.contains(sw)) return false;
        remove(p);
        nodes[p].box = sw.expand(margin);
        insert(p);
        return true;
    }
    template <class Fn> void query(const AABB<T>& b, Fn&& fn) const {  // aabb_tree.h:101-119
        if (root < 0) return;
        std::vector<int32_t> stack;
        stack.reserve(64);
        stack.push_back(root);
        while (!stack.empty()) {
            int32_t i = stack.back();
            stack.pop_back();
            const Node& n = nodes[i];
            if (!n.box.overlaps(b)) continue;
            if (n.c1 < 0) fn(n.data);
            else { stack.push_back(n.c1); stack.push_back(n.c2); }
        }
    }
private:
    int32_t alloc() {
        if (free_list < 0) { nodes.push_back(Node{}); return int32_t(nodes.size()) - 1; }
        int32_t i = free_list;
        free_list = nodes[i].parent;
        nodes[i] = Node{};
        return i;
    }
    void free_node(int32_t i) { nodes[i].height = -1; nodes[i].parent = free_list; free_list = i; }
    T child_cost(int32_t c, const AABB<T>& leaf) const {
        AABB<T> comb = AABB<T>::merge(nodes[c].box, leaf);
        if (nodes[c].c1 < 0) return comb.area();
        return comb.area() - nodes[c].box.area();
    }
    void insert(int32_t leaf) {  // aabb_tree.h:218-260
        if (root < 0) { root = leaf; nodes[leaf].parent = -1; return; }
        AABB<T> lb = nodes[leaf].box;
        int32_t i = root;
        while (nodes[i].c1 >= 0) {
            int32_t c1 = nodes[i].c1, c2 = nodes[i].c2;
            T area = nodes[i].box.area();
            T comb = AABB<T>::merge(nodes[i].box, lb).area();
            T cost = T(2) * comb;
            T inherit = T(2) * (comb - area);
            T cost1 = child_cost(c1, lb) + inherit;
            T cost2 = child_cost(c2, lb) + inherit;
            if (cost < cost1 && cost < cost2) break;
            i = (cost1 < cost2) ? c1 : c2;
        }
        int32_t sib = i, old_parent = nodes[sib].parent, np = alloc();
        nodes[np].parent = old_parent;
        nodes[np].box = AABB<T>::merge(lb, nodes[sib].box);
        nodes[np].height = nodes[sib].height + 1;
        nodes[np].c1 = sib;
        nodes[np].c2 = leaf;
        nodes[sib].parent = np;
        nodes[leaf].parent = np;
        if (old_parent < 0) root = np;
        else if (nodes[old_parent].c1 == sib) nodes[old_parent].c1 = np;
        else nodes[old_parent].c2 = np;
        refit(nodes[leaf].parent);
    }
    void remove(int32_t leaf) {  // aabb_tree.h:262-283
        if (leaf == root) { root = -1; return; }
        int32_t parent = nodes[leaf].parent, gp = nodes[parent].parent;
        int32_t sib = nodes[parent].c1 == leaf ? nodes[parent].c2 : nodes[parent].c1;
        if (gp < 0) { root = sib; nodes[sib].parent = -1; }
        else {
            if (nodes[gp].c1 == parent) nodes[gp].c1 = sib; else nodes[gp].c2 = sib;
            nodes[sib].parent = gp;
            refit(gp);
        }
        free_node(parent);
    }
    void refit(int32_t i) {
        while (i >= 0) {
            int32_t c1 = nodes[i].c1, c2 = nodes[i].c2;
            nodes[i].box = AABB<T>::merge(nodes[c1].box, nodes[c2].box);
            nodes[i].height = 1 + std::max(nodes[c1].height, nodes[c2].height);
            i = nodes[i].parent;
        }
    }
};

}  // namespace pxo

#include "pxo_hull.hpp"
#include "pxo_mesh.hpp"
