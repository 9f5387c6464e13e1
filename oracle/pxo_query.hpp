// ORACLE -- test infrastructure only. Nothing in the product may include, link or run this.
//
// Scene raycasts restated from the reference: Ray::intersect (geometry/ray.h:47-60, the slab test
// the AABB trees cull with), ray_vs_sphere (geometry/ray.h:62-78), ray_vs_obb (:101-125),
// ray_vs_capsule (:138-194), raycast_shape's sphere/box/capsule dispatch (query/queries.h:120-134),
// QueryFilter::accepts without the predicate (query/queries.h:92-98) and the world queries
// raycast / raycast_any / raycast_all (world.h:766-847). The trees visit every proxy whose fat
// AABB the ray's slab test accepts (a leaf's box lies inside its ancestors', and the slab test is
// monotonic in the box), so the world queries here test every live shape against its fat AABB and
// then exactly. Equal distances are broken by the lower shape slot (the reference's tie order is
// its tree traversal order, which nothing pins).
#pragma once
#include "pxo_step.hpp"

#include <algorithm>
#include <limits>
#include <vector>

namespace pxo {

template <class T> struct Ray {
    V3<T> o, d;
    T max_d = std::numeric_limits<T>::max();
};
struct QueryFilter {
    uint32_t layer_mask = ~0u;
    bool include_triggers = true;
    uint32_t ignore = ~0u;  // body index, ~0u: none
};
template <class T> struct RayHit {
    V3<T> point, normal;
    T distance = T(0);
    uint32_t body = ~0u, shape = ~0u;
};

// std::max / std::min over an initializer list: the first largest / smallest element
template <class T> inline T max4(T a, T b, T c, T d) {
    T m = a;
    if (m < b) m = b;
    if (m < c) m = c;
    if (m < d) m = d;
    return m;
}
template <class T> inline T min4(T a, T b, T c, T d) {
    T m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    if (d < m) m = d;
    return m;
}

template <class T> bool ray_aabb(const Ray<T>& ray, const V3<T>& mn, const V3<T>& mx) {  // ray.h:47-60
    const T e = T(1e-9f);
    V3<T> inv(T(1) / (ray.d.x + e), T(1) / (ray.d.y + e), T(1) / (ray.d.z + e));
    V3<T> t0 = (mn - ray.o) * inv, t1 = (mx - ray.o) * inv;
    V3<T> tmin = vmin(t0, t1), tmax = vmax(t0, t1);
    T t_near = max4(tmin.x, tmin.y, tmin.z, T(0));
    T t_far = min4(tmax.x, tmax.y, tmax.z, ray.max_d);
    return !(t_far < t_near);
}

template <class T> bool ray_sphere(const Ray<T>& ray, const V3<T>& c, T r, T& t, V3<T>& n) {  // ray.h:62-78
    V3<T> oc = ray.o - c;
    T b = dot(oc, ray.d);
    T cc = dot(oc, oc) - r * r;
    T disc = b * b - cc;
    if (disc < T(0)) return false;
    T s = std::sqrt(disc);
    T t0 = -b - s, t1 = -b + s;
    T hit = t0 >= T(0) ? t0 : t1;
    if (hit < T(0) || hit > ray.max_d) return false;
    t = hit;
    n = normalize((ray.o + ray.d * t) - c);
    return true;
}

template <class T> bool ray_obb(const Ray<T>& ray, const OBB<T>& box, T& t, V3<T>& n) {  // ray.h:101-125
    Q<T> inv = qinverse(box.q);
    V3<T> lo = qrot(inv, ray.o - box.c);
    V3<T> ld = qrot(inv, ray.d);
    const T e = T(1e-9f);
    V3<T> invd(T(1) / (ld.x + e), T(1) / (ld.y + e), T(1) / (ld.z + e));
    V3<T> t0 = (-box.he - lo) * invd, t1 = (box.he - lo) * invd;
    V3<T> tmin = vmin(t0, t1), tmax = vmax(t0, t1);
    T t_near = max4(tmin.x, tmin.y, tmin.z, T(0));
    T t_far = min4(tmax.x, tmax.y, tmax.z, ray.max_d);
    if (t_far < t_near) return false;
    int axis = 0;
    if (tmin.y > tmin.x) axis = 1;
    T ta = axis == 0 ? tmin.x : tmin.y;
    if (tmin.z > ta) axis = 2;
    T lda = axis == 0 ? ld.x : (axis == 1 ? ld.y : ld.z);
    V3<T> ln(T(0), T(0), T(0));
    T sgn = lda >= T(0) ? T(-1) : T(1);
    if (axis == 0) ln.x = sgn; else if (axis == 1) ln.y = sgn; else ln.z = sgn;
    t = t_near;
    n = qrot(box.q, ln);
    return true;
}

template <class T> bool ray_capsule(const Ray<T>& ray, const Cap<T>& cap, T& t, V3<T>& n) {  // ray.h:138-194
    V3<T> axis = cap.b - cap.a;
    T axis_len2 = dot(axis, axis);
    bool found = false;
    T best_t = ray.max_d;
    V3<T> best_n(T(0), T(0), T(0));
    if (axis_len2 > k_epsilon<T>()) {
        V3<T> axis_n = axis / std::sqrt(axis_len2);
        V3<T> oc = ray.o - cap.a;
        V3<T> d_perp = ray.d - axis_n * dot(ray.d, axis_n);
        V3<T> oc_perp = oc - axis_n * dot(oc, axis_n);
        T a = dot(d_perp, d_perp);
        if (a > k_epsilon<T>()) {
            T b = dot(d_perp, oc_perp);
            T c = dot(oc_perp, oc_perp) - cap.r * cap.r;
            T disc = b * b - a * c;
            if (disc >= T(0)) {
                T s = std::sqrt(disc);
                T cands[2] = {(-b - s) / a, (-b + s) / a};
                for (int k = 0; k < 2; ++k) {
                    T cand = cands[k];
                    if (cand < T(0) || cand >= best_t) continue;
                    T along = dot((ray.o + ray.d * cand) - cap.a, axis_n);
                    if (along < T(0) || along > std::sqrt(axis_len2)) continue;
                    V3<T> hp = ray.o + ray.d * cand;
                    V3<T> cl = cap.a + axis_n * along;
                    best_t = cand;
                    best_n = normalize(hp - cl);
                    found = true;
                    break;
                }
            }
        }
    }
    const V3<T> ends[2] = {cap.a, cap.b};
    for (int k = 0; k < 2; ++k) {
        Ray<T> capped = ray;
        capped.max_d = best_t;
        T ct;
        V3<T> cn;
        if (ray_sphere(capped, ends[k], cap.r, ct, cn)) {
            best_t = ct;
            best_n = cn;
            found = true;
        }
    }
    if (!found) return false;
    t = best_t;
    n = best_n;
    return true;
}

template <class T> bool raycast_shape(const Ray<T>& ray, const Shape<T>& s, const V3<T>& pos, const Q<T>& rot, T& t,
                                      V3<T>& n) {  // query/queries.h:120-134
    if (s.type == PXG_SPHERE) {
        Sph<T> sp = world_sphere(s, pos, rot);
        return ray_sphere(ray, sp.c, sp.r, t, n);
    }
    if (s.type == PXG_BOX) return ray_obb(ray, world_obb(s, pos, rot), t, n);
    if (s.type == PXG_CAPSULE) return ray_capsule(ray, world_capsule(s, pos, rot), t, n);
    return false;
}

// Every live shape the filter accepts and the ray hits within max distance, as (distance, shape)
// with the hit; `closest_only` keeps the nearest (lower slot on ties).
template <class T> std::vector<RayHit<T>> world_raycast(World<T>& w, const Ray<T>& ray, const QueryFilter& f,
                                                        bool closest_only) {
    std::vector<RayHit<T>> out;
    for (uint32_t i = 0; i < w.shapes.size(); ++i) {
        const Shape<T>& s = w.shapes[i];
        if (i >= w.proxy.size() || w.proxy[i] < 0 || s.owner >= w.bodies.size() || !w.bodies[s.owner].alive) continue;
        if (!((f.layer_mask >> (s.layer & 31u)) & 1u)) continue;
        if (!f.include_triggers && s.trigger) continue;
        if (f.ignore != ~0u && s.owner == f.ignore) continue;
        const AABB<T> box = w.fat(i);
        if (!ray_aabb(ray, box.mn, box.mx)) continue;
        const Body<T>& b = w.bodies[s.owner];
        T t;
        V3<T> n;
        if (!raycast_shape(ray, s, b.pos, b.q, t, n)) continue;
        RayHit<T> h;
        h.point = ray.o + ray.d * t;
        h.normal = n;
        h.distance = t;
        h.body = s.owner;
        h.shape = i;
        if (closest_only) {
            if (out.empty()) out.push_back(h);
            else if (t < out[0].distance) out[0] = h;
        } else {
            out.push_back(h);
        }
    }
    if (!closest_only)
        std::sort(out.begin(), out.end(), [](const RayHit<T>& a, const RayHit<T>& b) {
            return a.distance < b.distance || (a.distance == b.distance && a.shape < b.shape);
        });
    return out;
}

// overlap_sphere / overlap_box / overlap_capsule (world.h:1002-1060): the query AABB against the
// fat AABBs (AABB::overlaps, inclusive, geometry/aabb.h:35-39; the trees' query), the filter,
// then shape_overlaps_sphere / _obb / _capsule (query/queries.h:139-258). Bodies in slot order
// (the reference lists them in tree order).
template <class T> struct OverlapQuery {
    int type = PXG_SPHERE;
    Sph<T> sph;
    OBB<T> box;
    Cap<T> cap;
};
template <class T> bool shape_overlaps(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot, const OverlapQuery<T>& q) {
    V3<T> n, p;
    T pen;
    Manifold<T> m;
    if (q.type == PXG_SPHERE) {
        if (s.type == PXG_SPHERE) return sphere_vs_sphere(world_sphere(s, pos, rot), q.sph, n, pen, p);
        if (s.type == PXG_BOX) return sphere_vs_box(q.sph, world_obb(s, pos, rot), n, pen, p);
        return capsule_vs_sphere(world_capsule(s, pos, rot), q.sph, n, pen, p);
    }
    if (q.type == PXG_BOX) {
        if (s.type == PXG_SPHERE) return sphere_vs_box(world_sphere(s, pos, rot), q.box, n, pen, p);
        if (s.type == PXG_BOX) return box_box(world_obb(s, pos, rot), q.box, m, false);
        return capsule_vs_box(world_capsule(s, pos, rot), q.box, m);
    }
    if (s.type == PXG_SPHERE) return capsule_vs_sphere(q.cap, world_sphere(s, pos, rot), n, pen, p);
    if (s.type == PXG_BOX) return capsule_vs_box(q.cap, world_obb(s, pos, rot), m);
    return capsule_vs_capsule(q.cap, world_capsule(s, pos, rot), n, pen, p);
}
template <class T> AABB<T> overlap_bounds(const OverlapQuery<T>& q) {
    AABB<T> b;
    if (q.type == PXG_SPHERE) {
        V3<T> r(q.sph.r, q.sph.r, q.sph.r);
        b.mn = q.sph.c - r;
        b.mx = q.sph.c + r;
    } else if (q.type == PXG_BOX) {
        b = obb_bounds(q.box);
    } else {
        V3<T> r(q.cap.r, q.cap.r, q.cap.r);
        b.mn = vmin(q.cap.a, q.cap.b) - r;
        b.mx = vmax(q.cap.a, q.cap.b) + r;
    }
    return b;
}
template <class T> std::vector<uint32_t> world_overlap(World<T>& w, const OverlapQuery<T>& q, const QueryFilter& f) {
    std::vector<uint32_t> out;
    const AABB<T> qb = overlap_bounds(q);
    for (uint32_t i = 0; i < w.shapes.size(); ++i) {
        const Shape<T>& s = w.shapes[i];
        if (i >= w.proxy.size() || w.proxy[i] < 0 || s.owner >= w.bodies.size() || !w.bodies[s.owner].alive) continue;
        if (!((f.layer_mask >> (s.layer & 31u)) & 1u)) continue;
        if (!f.include_triggers && s.trigger) continue;
        if (f.ignore != ~0u && s.owner == f.ignore) continue;
        if (!w.fat(i).overlaps(qb)) continue;
        const Body<T>& b = w.bodies[s.owner];
        if (shape_overlaps(s, b.pos, b.q, q)) out.push_back(s.owner);
    }
    return out;
}

// ---------------- shape casts (query/sweep.h, world.h:877-950) ----------------
constexpr float k_sweep_tolerance = 1e-4f;  // sweep.h:68

template <class T> struct CastGeom {  // sweep.h:86-121: core segment (a == b for a sphere) + radius, or the box
    int type = PXG_SPHERE;
    V3<T> a, b;
    T r = T(0);
    OBB<T> box;
};
template <class T> CastGeom<T> cast_geometry(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot) {
    CastGeom<T> g;
    g.type = s.type;
    if (s.type == PXG_SPHERE) {
        Sph<T> sp = world_sphere(s, pos, rot);
        g.a = g.b = sp.c;
        g.r = sp.r;
    } else if (s.type == PXG_CAPSULE) {
        Cap<T> c = world_capsule(s, pos, rot);
        g.a = c.a; g.b = c.b; g.r = c.r;
    } else {
        g.box = world_obb(s, pos, rot);
    }
    return g;
}

template <class T> inline T vc(const V3<T>& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
template <class T> inline void vset(V3<T>& v, int i, T x) { if (i == 0) v.x = x; else if (i == 1) v.y = x; else v.z = x; }

// closest_points_segment_box (collision/distance.h:41-102): exact, over the pieces between the
// segment's slab-plane crossings; returns the squared distance
template <class T> T closest_seg_box(const V3<T>& p0, const V3<T>& p1, const OBB<T>& box, V3<T>& on_seg, V3<T>& on_box) {
    V3<T> ax[3];
    obb_axes(box, ax);
    const V3<T> rel0 = p0 - box.c, rel1 = p1 - box.c;
    V3<T> o, d;
    for (int i = 0; i < 3; ++i) {
        vset(o, i, dot(rel0, ax[i]));
        vset(d, i, dot(rel1, ax[i]) - vc(o, i));
    }
    const V3<T>& h = box.he;
    T breaks[8];
    int count = 0;
    breaks[count++] = T(0);
    breaks[count++] = T(1);
    for (int i = 0; i < 3; ++i) {
        if (std::abs(vc(d, i)) < T(1e-12f)) continue;
        const T planes[2] = {-vc(h, i), vc(h, i)};
        for (int k = 0; k < 2; ++k) {
            T sv = (planes[k] - vc(o, i)) / vc(d, i);
            if (sv > T(0) && sv < T(1)) breaks[count++] = sv;
        }
    }
    std::sort(breaks, breaks + count);
    const V3<T> nh = -h;
    auto sq_at = [&](T sv) {
        V3<T> p = o + d * sv;
        V3<T> ex = p - V3<T>(gclamp(p.x, nh.x, h.x), gclamp(p.y, nh.y, h.y), gclamp(p.z, nh.z, h.z));
        return dot(ex, ex);
    };
    T best_s = T(0);
    T best = sq_at(T(0));
    for (int k = 0; k + 1 < count; ++k) {
        T s0 = breaks[k], s1 = breaks[k + 1];
        if (s1 - s0 < T(1e-9f)) continue;
        V3<T> mid = o + d * (T(0.5f) * (s0 + s1));
        T a = T(0), b = T(0);
        for (int i = 0; i < 3; ++i) {
            T c;
            if (vc(mid, i) > vc(h, i)) c = vc(h, i);
            else if (vc(mid, i) < -vc(h, i)) c = -vc(h, i);
            else continue;
            a += vc(d, i) * vc(d, i);
            b += vc(d, i) * (vc(o, i) - c);
        }
        T sv = a > T(1e-12f) ? sclamp(-b / a, s0, s1) : s0;
        T v = sq_at(sv);
        if (v < best) { best = v; best_s = sv; }
    }
    T end_v = sq_at(T(1));
    if (end_v < best) { best = end_v; best_s = T(1); }
    V3<T> local = o + d * best_s;
    V3<T> cl(gclamp(local.x, nh.x, h.x), gclamp(local.y, nh.y, h.y), gclamp(local.z, nh.z, h.z));
    on_seg = p0 + (p1 - p0) * best_s;
    on_box = box.c + ax[0] * cl.x + ax[1] * cl.y + ax[2] * cl.z;
    return best;
}

// push_axis_, project_obb_, segment_box_mtv_ (sweep.h:215-230, 304-338)
template <class T> inline void push_axis(V3<T>* axes, int& count, const V3<T>& L) {
    T len2 = dot(L, L);
    if (len2 < T(1e-10f)) return;
    axes[count++] = L / std::sqrt(len2);
}
template <class T> inline void project_obb(const OBB<T>& box, const V3<T> bx[3], const V3<T>& L, T& lo, T& hi) {
    T c = dot(box.c, L);
    T r = box.he.x * std::abs(dot(bx[0], L)) + box.he.y * std::abs(dot(bx[1], L)) + box.he.z * std::abs(dot(bx[2], L));
    lo = c - r;
    hi = c + r;
}
template <class T> void segment_box_mtv(const V3<T>& p0, const V3<T>& p1, T radius, const OBB<T>& box, V3<T>& normal,
                                        T& depth) {
    V3<T> bx[3];
    obb_axes(box, bx);
    V3<T> axes[6];
    int count = 0;
    for (int i = 0; i < 3; ++i) push_axis(axes, count, bx[i]);
    for (int i = 0; i < 3; ++i) push_axis(axes, count, cross(p1 - p0, bx[i]));
    depth = std::numeric_limits<T>::max();
    normal = V3<T>(T(0), T(0), T(1));
    for (int k = 0; k < count; ++k) {
        const V3<T>& L = axes[k];
        T s0 = dot(p0, L), s1 = dot(p1, L);
        T b0, b1;
        project_obb(box, bx, L, b0, b1);
        T lo = gmin(s0, s1), hi = gmax(s0, s1);
        T overlap = gmin(hi - b0, b1 - lo) + radius;
        if (overlap < depth) {
            depth = overlap;
            normal = (lo + hi) >= (b0 + b1) ? L : -L;
        }
    }
}

// convex_penetration for a sphere or capsule probe (sweep.h:339-384)
template <class T> bool convex_penetration(const Shape<T>& probe, const V3<T>& pos, const Q<T>& rot, const Shape<T>& target,
                                           const V3<T>& tpos, const Q<T>& trot, V3<T>& normal, T& depth, V3<T>& point) {
    const bool probe_core_vs_box = probe.type == PXG_CAPSULE && target.type == PXG_BOX;
    const bool box_vs_target_core = probe.type == PXG_BOX && target.type == PXG_CAPSULE;
    if (probe_core_vs_box || box_vs_target_core) {
        CastGeom<T> core = probe_core_vs_box ? cast_geometry(probe, pos, rot) : cast_geometry(target, tpos, trot);
        OBB<T> box = probe_core_vs_box ? world_obb(target, tpos, trot) : world_obb(probe, pos, rot);
        V3<T> on_seg, on_box;
        T dist = std::sqrt(closest_seg_box(core.a, core.b, box, on_seg, on_box));
        if (dist >= core.r) return false;
        if (dist > k_epsilon<T>()) {
            normal = (on_seg - on_box) / dist;
            depth = core.r - dist;
        } else {
            segment_box_mtv(core.a, core.b, core.r, box, normal, depth);
        }
        point = on_box;
        if (!probe_core_vs_box) normal = -normal;  // the probe is the box: push it away from the core
        return depth > T(0);
    }
    Manifold<T> m;
    if (!generate_contacts(0u, probe, pos, rot, 0u, target, tpos, trot, m, false)) return false;
    normal = -m.n;
    depth = T(0);
    V3<T> sum(T(0), T(0), T(0));
    for (int k = 0; k < m.count; ++k) {
        depth = gmax(depth, m.p[k].pen);
        sum += m.p[k].pos;
    }
    point = m.count > 0 ? sum / T(m.count) : tpos;
    return depth > T(0);
}

// obb_corners_, swept_sat_, sweep_box_box_, polytope_contact_point_ (sweep.h:128-140, 175-258, 282-295)
template <class T> void obb_corners(const OBB<T>& box, V3<T> out[8]) {
    V3<T> ax[3];
    obb_axes(box, ax);
    int k = 0;
    for (int sx = -1; sx <= 1; sx += 2)
        for (int sy = -1; sy <= 1; sy += 2)
            for (int sz = -1; sz <= 1; sz += 2)
                out[k++] = box.c + ax[0] * (T(sx) * box.he.x) + ax[1] * (T(sy) * box.he.y) + ax[2] * (T(sz) * box.he.z);
}
enum class Advance { Hit, Miss, AtStart };
template <class T> Advance sweep_box_box(const OBB<T>& a, const OBB<T>& b, const V3<T>& dir, T maxd, T& toi, V3<T>& axis,
                                         V3<T>& mtv_axis, T& mtv_depth) {
    V3<T> ax[3], bx[3];
    obb_axes(a, ax);
    obb_axes(b, bx);
    V3<T> axes[15];
    int count = 0;
    for (int i = 0; i < 3; ++i) push_axis(axes, count, ax[i]);
    for (int i = 0; i < 3; ++i) push_axis(axes, count, bx[i]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) push_axis(axes, count, cross(ax[i], bx[j]));
    T t_first = -std::numeric_limits<T>::max(), t_last = std::numeric_limits<T>::max();
    mtv_depth = std::numeric_limits<T>::max();
    for (int k = 0; k < count; ++k) {
        const V3<T>& L = axes[k];
        T a0, a1, b0, b1;
        project_obb(a, ax, L, a0, a1);
        project_obb(b, bx, L, b0, b1);
        T v = dot(dir, L);
        T overlap = gmin(a1 - b0, b1 - a0);
        if (overlap < mtv_depth) {
            mtv_depth = overlap;
            mtv_axis = (a0 + a1) >= (b0 + b1) ? L : -L;
        }
        if (std::abs(v) < T(1e-7f)) {
            if (overlap <= T(k_sweep_tolerance)) return Advance::Miss;
            continue;
        }
        T enter = v > T(0) ? (b0 - a1) / v : (b1 - a0) / v;
        T exit = v > T(0) ? (b1 - a0) / v : (b0 - a1) / v;
        if (enter > t_first) {
            t_first = enter;
            axis = v > T(0) ? -L : L;
        }
        t_last = gmin(t_last, exit);
        if (t_first > t_last || t_first > maxd || t_last < T(0)) return Advance::Miss;
    }
    if (t_first <= T(k_sweep_tolerance)) return Advance::AtStart;
    toi = t_first;
    return Advance::Hit;
}
template <class T> V3<T> polytope_contact_point(const OBB<T>& box_at_toi, const V3<T>& n, const OBB<T>& target) {
    V3<T> corners[8];
    obb_corners(box_at_toi, corners);
    T lo = std::numeric_limits<T>::max();
    for (int k = 0; k < 8; ++k) lo = gmin(lo, dot(corners[k], n));
    V3<T> sum(T(0), T(0), T(0));
    int count = 0;
    T eps = T(1e-3f) * max4(box_at_toi.he.x, box_at_toi.he.y, box_at_toi.he.z, T(1e-3f));
    for (int k = 0; k < 8; ++k) {
        if (dot(corners[k], n) <= lo + eps) {
            sum += corners[k];
            ++count;
        }
    }
    return closest_point_on_obb(sum / T(count > 1 ? count : 1), target);
}

template <class T> struct SweepHit {
    T distance = T(0);
    V3<T> normal{T(0), T(0), T(1)}, point{T(0), T(0), T(0)};
    bool started_inside = false;
};

// sweep_convex (sweep.h:149-168 advance_, 408-416 resolve_at_start_, 422-505) for a sphere,
// capsule or box cast shape against a sphere, box or capsule target
template <class T> bool sweep_convex(const Shape<T>& shape, const V3<T>& pos, const Q<T>& rot, const V3<T>& dir, T maxd,
                                     const Shape<T>& target, const V3<T>& tpos, const Q<T>& trot, SweepHit<T>& out) {
    const CastGeom<T> cast = cast_geometry(shape, pos, rot);
    T toi = T(0);
    V3<T> n(T(0), T(0), T(1)), p(T(0), T(0), T(0)), mtv_axis(T(0), T(0), T(1));
    T mtv_depth = T(0);
    const bool polytopes = shape.type == PXG_BOX && target.type == PXG_BOX;
    Advance result;
    if (polytopes) {
        OBB<T> tb = world_obb(target, tpos, trot);
        result = sweep_box_box(cast.box, tb, dir, maxd, toi, n, mtv_axis, mtv_depth);
        if (result == Advance::Hit) {
            OBB<T> moved = cast.box;
            moved.c += dir * toi;
            p = polytope_contact_point(moved, n, tb);
        } else if (result == Advance::AtStart) {
            p = closest_point_on_obb(cast.box.c, tb);
        }
    } else {
        const bool box_cast = shape.type == PXG_BOX;
        auto dist_at = [&](T t, V3<T>& normal, V3<T>& point) {
            if (box_cast) {  // the moving box against the target's core
                CastGeom<T> core = cast_geometry(target, tpos, trot);
                OBB<T> moved = cast.box;
                moved.c += dir * t;
                V3<T> on_seg, on_box;
                T dist = std::sqrt(closest_seg_box(core.a, core.b, moved, on_seg, on_box));
                if (dist > k_epsilon<T>()) normal = (on_box - on_seg) / dist;
                point = on_seg + normal * core.r;
                return dist - core.r;
            }
            V3<T> a = cast.a + dir * t, b = cast.b + dir * t;
            V3<T> on_cast, on_target;
            T target_radius = T(0);
            if (target.type == PXG_BOX) {
                closest_seg_box(a, b, world_obb(target, tpos, trot), on_cast, on_target);
            } else {
                CastGeom<T> core = cast_geometry(target, tpos, trot);
                closest_points_seg_seg(a, b, core.a, core.b, on_cast, on_target);
                target_radius = core.r;
            }
            V3<T> diff = on_cast - on_target;
            T dist = length(diff);
            if (dist > k_epsilon<T>()) normal = diff / dist;
            point = on_target + normal * target_radius;
            return dist - cast.r - target_radius;
        };
        T t = T(0);
        T dd = dist_at(t, n, p);
        if (dd <= T(k_sweep_tolerance)) {
            result = Advance::AtStart;
        } else {
            result = Advance::Hit;
            for (int iter = 0; iter < 64; ++iter) {
                T closing = -dot(dir, n);
                if (closing <= T(1e-6f)) { result = Advance::Miss; break; }
                t += dd / closing;
                if (t > maxd) { result = Advance::Miss; break; }
                dd = dist_at(t, n, p);
                if (dd <= T(k_sweep_tolerance)) break;
            }
            toi = t;
        }
    }
    if (result == Advance::Miss) return false;
    if (result == Advance::AtStart) {
        V3<T> normal = polytopes ? mtv_axis : n, point = p, pen_n, pen_p;
        T depth = polytopes ? mtv_depth : T(0), pen_d = T(0);
        if (!polytopes && convex_penetration(shape, pos, rot, target, tpos, trot, pen_n, pen_d, pen_p)) {
            normal = pen_n;
            point = pen_p;
            depth = pen_d;
        }
        if (dot(dir, normal) >= T(-1e-4f)) return false;  // resolve_at_start_
        out.distance = T(0);
        out.normal = normal;
        out.point = point;
        out.started_inside = depth > T(k_sweep_tolerance);
        return true;
    }
    out.distance = toi;
    out.normal = n;
    out.point = p;
    out.started_inside = false;
    return true;
}

// shape_cast (world.h:877-909): the swept AABB against the fat AABBs, the
// filter, then each candidate's sweep; the nearest wins (lower slot on ties: the reference's tie
// order is its tree order)
template <class T> bool world_shape_cast(World<T>& w, const Shape<T>& shape, const V3<T>& pos, const Q<T>& rot,
                                         const V3<T>& dir, T maxd, const QueryFilter& f, RayHit<T>& hit, bool& started) {
    if (!(maxd >= T(0))) return false;
    T dir_len = length(dir);
    if (dir_len < T(1e-8f)) return false;
    const V3<T> ud = dir / dir_len;
    AABB<T> b0 = world_bounds(shape, pos, rot), b1 = world_bounds(shape, pos + ud * maxd, rot);
    AABB<T> swept;
    swept.mn = vmin(b0.mn, b1.mn);
    swept.mx = vmax(b0.mx, b1.mx);
    bool found = false;
    T best = maxd;
    for (uint32_t i = 0; i < w.shapes.size(); ++i) {
        const Shape<T>& s = w.shapes[i];
        if (i >= w.proxy.size() || w.proxy[i] < 0 || s.owner >= w.bodies.size() || !w.bodies[s.owner].alive) continue;
        if (!((f.layer_mask >> (s.layer & 31u)) & 1u)) continue;
        if (!f.include_triggers && s.trigger) continue;
        if (f.ignore != ~0u && s.owner == f.ignore) continue;
        if (!w.fat(i).overlaps(swept)) continue;
        if (found && best <= T(0)) continue;
        const Body<T>& b = w.bodies[s.owner];
        SweepHit<T> r;
        if (!sweep_convex(shape, pos, rot, ud, best, s, b.pos, b.q, r)) continue;
        if (found && r.distance >= best) continue;
        best = r.distance;
        hit.point = r.point;
        hit.normal = r.normal;
        hit.distance = r.distance;
        hit.body = s.owner;
        hit.shape = i;
        started = r.started_inside;
        found = true;
    }
    return found;
}

// compute_penetration (world.h:963-995) for a sphere, box or capsule query shape: live shapes
// whose fat AABB meets the query shape's world bounds (query_aabb_, inclusive) and that pass the
// filter, each through convex_penetration; slot order (the reference's order is its tree order)
template <class T> struct Penetration {
    V3<T> normal, point;
    T depth = T(0);
    uint32_t body = 0, shape = 0;
};
template <class T> std::vector<Penetration<T>> world_penetration(World<T>& w, const Shape<T>& shape, const V3<T>& pos,
                                                                 const Q<T>& rot, const QueryFilter& f) {
    std::vector<Penetration<T>> out;
    const AABB<T> qb = world_bounds(shape, pos, rot);
    for (uint32_t i = 0; i < w.shapes.size(); ++i) {
        const Shape<T>& s = w.shapes[i];
        if (i >= w.proxy.size() || w.proxy[i] < 0 || s.owner >= w.bodies.size() || !w.bodies[s.owner].alive) continue;
        if (!((f.layer_mask >> (s.layer & 31u)) & 1u)) continue;
        if (!f.include_triggers && s.trigger) continue;
        if (f.ignore != ~0u && s.owner == f.ignore) continue;
        if (!w.fat(i).overlaps(qb)) continue;
        const Body<T>& b = w.bodies[s.owner];
        Penetration<T> p;
        if (!convex_penetration(shape, pos, rot, s, b.pos, b.q, p.normal, p.depth, p.point)) continue;
        p.body = s.owner;
        p.shape = i;
        out.push_back(p);
    }
    return out;
}

}  // namespace pxo
