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

}  // namespace pxo
