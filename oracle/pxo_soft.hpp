// ORACLE -- test infrastructure only. Nothing in the product may include, link or run this.
//
// XPBD cloth and hair, restated function by function: the cloth records (cloth/cloth.h), the
// grid builder, pins and tethers (cloth/cloth_builder.h), the particle-against-shape projection
// (cloth/cloth_collision.h, including the mesh's swept crossing test and proximity query and the
// hull's sphere probe), the cloth substep loop (cloth/cloth_solver.h: aerodynamics, integration,
// anchors, stretch and bend batches, tethers, self-collision, collisions, velocities and
// friction), and the strand records, builder and substep loop (hair/strand.h,
// hair/strand_builder.h, hair/strand_solver.h). The world's drivers (world.h:1749-1957) are
// World::step_cloths / step_strands in pxo_step.hpp.
//
// Transcendental functions (sin, cos, pow, acos) follow the world's Atan2Mode: Libm calls the
// float functions the reference calls; Rounded evaluates them in double and rounds once.
#pragma once
#include "pxo_world.hpp"
#include "pxo_joints.hpp"

#include <cmath>
#include <queue>
#include <vector>

namespace pxo {

template <class T> inline T t_sin(T x, Atan2Mode m) { return m == Atan2Mode::Rounded ? T(std::sin(double(x))) : std::sin(x); }
template <class T> inline T t_cos(T x, Atan2Mode m) { return m == Atan2Mode::Rounded ? T(std::cos(double(x))) : std::cos(x); }
template <class T> inline T t_acos(T x, Atan2Mode m) { return m == Atan2Mode::Rounded ? T(std::acos(double(x))) : std::acos(x); }
template <class T> inline T t_pow(T x, T y, Atan2Mode m) {
    return m == Atan2Mode::Rounded ? T(std::pow(double(x), double(y))) : std::pow(x, y);
}
// glm::mix(x, y, a) = x * (1 - a) + y * a
template <class T> inline V3<T> vmix(const V3<T>& x, const V3<T>& y, T a) { return x * (T(1) - a) + y * a; }
template <class T> inline T smix(T x, T y, T a) { return x * (T(1) - a) + y * a; }
// util::safe_normalize (util/math.h:46-50)
template <class T> inline V3<T> safe_normalize(const V3<T>& v) {
    T l2 = dot(v, v);
    if (l2 < k_epsilon<T>() * k_epsilon<T>()) return V3<T>(T(0));
    return v * (T(1) / std::sqrt(l2));
}

// ---- records (cloth/cloth.h:46-245) ----

template <class T> struct CParticle {
    V3<T> pos, prev, vel;
    T inv_mass = T(1);
};
template <class T> struct CConstraint {
    uint32_t a = 0, b = 0;
    T rest = T(0), lambda = T(0);
};
struct CBatch { uint32_t begin = 0, end = 0; };
template <class T> struct CTether {
    uint32_t particle = 0, anchor = 0;
    T max_len = T(0);
};
template <class T> struct CAnchor {  // a body index with generation, or none: a static pin
    uint32_t particle = 0;
    bool has_body = false;
    uint32_t body = 0, generation = 0;
    V3<T> local, world, prev_world;
};
template <class T> struct ClothParams {
    T stretch_compliance = T(0), bend_compliance = T(5e-5f), damping = T(0.05f), thickness = T(0.03f);
    T friction = T(0.5f), gravity_scale = T(1), max_velocity = T(100);
    V3<T> external_acceleration, wind;
    T wind_turbulence = T(0), air_drag = T(0.02f), air_lift = T(0.02f);
    bool self_collision = false;
    T self_distance = T(0), tether_scale = T(1.02f);
    uint32_t substeps = 0, iterations = 0;
    T sleep_threshold = T(0.05f), sleep_time = T(0.5f);
    uint32_t layer = 0;
};
template <class T> struct Cloth {
    std::vector<CParticle<T>> particles;
    std::vector<CConstraint<T>> stretch, bend;
    std::vector<CBatch> stretch_batches, bend_batches;
    std::vector<CTether<T>> tethers;
    std::vector<CAnchor<T>> anchors;
    std::vector<uint32_t> triangles;
    uint32_t columns = 0, rows = 0;
    ClothParams<T> params;
    AABB<T> bounds;
    bool enabled = true, awake = true;
    T sleep_timer = T(0), time = T(0);
};
template <class T> AABB<T> particle_bounds(const std::vector<CParticle<T>>& ps) {  // cloth.h:320-327
    AABB<T> out;
    for (const CParticle<T>& p : ps) {
        out.mn = vmin(out.mn, p.pos);
        out.mx = vmax(out.mx, p.pos);
    }
    return out;
}
template <class T> T max_particle_speed(const std::vector<CParticle<T>>& ps) {  // cloth.h:330-336
    T m2 = T(0);
    for (const CParticle<T>& p : ps) m2 = gmax(m2, dot(p.vel, p.vel));
    return std::sqrt(m2);
}

// ---- builder (cloth/cloth_builder.h) ----

template <class T> struct GridClothDesc {
    uint32_t columns = 21, rows = 21;
    T width = T(2), height = T(2);
    V3<T> center;
    Q<T> orientation;
    T total_mass = T(1);
    bool shear = true;
    ClothParams<T> params;
};
// make_grid_cloth (cloth_builder.h:103-239): particles row-major, structural edges by parity,
// optional diagonals, two-apart bend edges in four phases, two triangles per quad
template <class T> Cloth<T> make_grid_cloth(const GridClothDesc<T>& d) {
    Cloth<T> c;
    const uint32_t cols = std::max<uint32_t>(2, d.columns), rows = std::max<uint32_t>(2, d.rows);
    c.columns = cols;
    c.rows = rows;
    c.params = d.params;
    const uint32_t count = cols * rows;
    const T pm = d.total_mass > k_epsilon<T>() ? d.total_mass / T(count) : T(0);
    const T im = pm > k_epsilon<T>() ? T(1) / pm : T(0);
    const T dx = d.width / T(cols - 1), dy = d.height / T(rows - 1);
    c.particles.resize(count);
    for (uint32_t y = 0; y < rows; ++y)
        for (uint32_t x = 0; x < cols; ++x) {
            const V3<T> local(T(x) * dx - d.width * T(0.5f), T(y) * dy - d.height * T(0.5f), T(0));
            CParticle<T>& p = c.particles[y * cols + x];
            p.pos = d.center + qrot(d.orientation, local);
            p.prev = p.pos;
            p.inv_mass = im;
        }
    auto at = [cols](uint32_t x, uint32_t y) { return y * cols + x; };
    auto emit = [&c](std::vector<CConstraint<T>>& out, uint32_t a, uint32_t b) {
        CConstraint<T> k;
        k.a = a;
        k.b = b;
        k.rest = length(c.particles[a].pos - c.particles[b].pos);
        out.push_back(k);
    };
    auto close = [](std::vector<CConstraint<T>>& list, std::vector<CBatch>& batches, uint32_t begin) {
        const uint32_t end = uint32_t(list.size());
        if (end > begin) batches.push_back(CBatch{begin, end});
    };
    for (uint32_t par = 0; par < 2; ++par) {
        const uint32_t b = uint32_t(c.stretch.size());
        for (uint32_t y = 0; y < rows; ++y)
            for (uint32_t x = par; x + 1 < cols; x += 2) emit(c.stretch, at(x, y), at(x + 1, y));
        close(c.stretch, c.stretch_batches, b);
    }
    for (uint32_t par = 0; par < 2; ++par) {
        const uint32_t b = uint32_t(c.stretch.size());
        for (uint32_t y = par; y + 1 < rows; y += 2)
            for (uint32_t x = 0; x < cols; ++x) emit(c.stretch, at(x, y), at(x, y + 1));
        close(c.stretch, c.stretch_batches, b);
    }
    if (d.shear) {
        for (uint32_t par = 0; par < 2; ++par) {
            const uint32_t b = uint32_t(c.stretch.size());
            for (uint32_t y = 0; y + 1 < rows; ++y)
                for (uint32_t x = par; x + 1 < cols; x += 2) emit(c.stretch, at(x, y), at(x + 1, y + 1));
            close(c.stretch, c.stretch_batches, b);
        }
        for (uint32_t par = 0; par < 2; ++par) {
            const uint32_t b = uint32_t(c.stretch.size());
            for (uint32_t y = 0; y + 1 < rows; ++y)
                for (uint32_t x = par; x + 1 < cols; x += 2) emit(c.stretch, at(x + 1, y), at(x, y + 1));
            close(c.stretch, c.stretch_batches, b);
        }
    }
    for (uint32_t ph = 0; ph < 4; ++ph) {
        const uint32_t b = uint32_t(c.bend.size());
        for (uint32_t y = 0; y < rows; ++y)
            for (uint32_t x = ph; x + 2 < cols; x += 4) emit(c.bend, at(x, y), at(x + 2, y));
        close(c.bend, c.bend_batches, b);
    }
    for (uint32_t ph = 0; ph < 4; ++ph) {
        const uint32_t b = uint32_t(c.bend.size());
        for (uint32_t y = ph; y + 2 < rows; y += 4)
            for (uint32_t x = 0; x < cols; ++x) emit(c.bend, at(x, y), at(x, y + 2));
        close(c.bend, c.bend_batches, b);
    }
    for (uint32_t y = 0; y + 1 < rows; ++y)
        for (uint32_t x = 0; x + 1 < cols; ++x) {
            const uint32_t i00 = at(x, y), i10 = at(x + 1, y), i01 = at(x, y + 1), i11 = at(x + 1, y + 1);
            c.triangles.insert(c.triangles.end(), {i00, i10, i11, i00, i11, i01});
        }
    if (c.params.self_distance <= T(0)) c.params.self_distance = T(0.6f) * std::min(dx, dy);
    c.bounds = particle_bounds(c.particles);
    return c;
}
// pin_static / pin_to_body (cloth_builder.h:265-331): every free particle within radius is
// pinned (inv_mass 0) to the world point, or to the body at its current pose
template <class T> uint32_t pin_particles(Cloth<T>& c, const V3<T>& point, T radius, bool to_body, uint32_t body,
                                          uint32_t generation, const V3<T>& body_pos, const Q<T>& body_rot) {
    const T r2 = radius * radius;
    const Q<T> inv = qinverse(body_rot);
    uint32_t pinned = 0;
    for (uint32_t i = 0; i < c.particles.size(); ++i) {
        CParticle<T>& p = c.particles[i];
        if (p.inv_mass == T(0)) continue;
        const V3<T> d = p.pos - point;
        if (dot(d, d) > r2) continue;
        p.inv_mass = T(0);
        p.vel = V3<T>(T(0));
        CAnchor<T> a;
        a.particle = i;
        a.has_body = to_body;
        a.body = body;
        a.generation = generation;
        a.local = to_body ? qrot(inv, p.pos - body_pos) : V3<T>(T(0));
        a.world = a.prev_world = p.pos;
        c.anchors.push_back(a);
        ++pinned;
    }
    return pinned;
}
// build_tethers (cloth_builder.h:344-392): Dijkstra over the stretch graph from every anchor;
// each free particle is held within tether_scale x its geodesic distance of the nearest one
template <class T> void build_tethers(Cloth<T>& c) {
    c.tethers.clear();
    if (c.anchors.empty() || c.particles.empty()) return;
    const size_t n = c.particles.size();
    std::vector<std::vector<uint32_t>> adj(n);
    for (const CConstraint<T>& k : c.stretch) {
        adj[k.a].push_back(k.b);
        adj[k.b].push_back(k.a);
    }
    const T inf = std::numeric_limits<T>::max();
    std::vector<T> dist(n, inf);
    std::vector<uint32_t> source(n, 0);
    using E = std::pair<T, uint32_t>;
    std::priority_queue<E, std::vector<E>, std::greater<E>> q;
    for (const CAnchor<T>& a : c.anchors) {
        dist[a.particle] = T(0);
        source[a.particle] = a.particle;
        q.push({T(0), a.particle});
    }
    while (!q.empty()) {
        const E top = q.top();
        q.pop();
        if (top.first > dist[top.second]) continue;
        for (uint32_t nx : adj[top.second]) {
            const T cand = top.first + length(c.particles[nx].pos - c.particles[top.second].pos);
            if (cand >= dist[nx]) continue;
            dist[nx] = cand;
            source[nx] = source[top.second];
            q.push({cand, nx});
        }
    }
    const T scale = std::max(T(1), c.params.tether_scale);
    for (uint32_t i = 0; i < n; ++i) {
        if (c.particles[i].inv_mass == T(0) || dist[i] >= inf) continue;
        c.tethers.push_back(CTether<T>{i, source[i], dist[i] * scale});
    }
}

// ---- particle against a shape (cloth/cloth_collision.h) ----

template <class T> struct Collider {
    const Shape<T>* shape = nullptr;
    V3<T> pos;
    Q<T> rot;
    AABB<T> bounds;
};

// Ray::intersect (geometry/ray.h:31-42): slab test with a 1e-9 nudge against zero directions
template <class T> bool ray_box(const V3<T>& o, const V3<T>& dir, T maxd, const AABB<T>& b, T& t) {
    const V3<T> nd = dir + V3<T>(T(1e-9f));
    const V3<T> inv(T(1) / nd.x, T(1) / nd.y, T(1) / nd.z);
    const V3<T> t0 = (b.mn - o) * inv, t1 = (b.mx - o) * inv;
    const V3<T> lo = vmin(t0, t1), hi = vmax(t0, t1);
    T tn = lo.x;  // std::max({...}): the first largest
    if (tn < lo.y) tn = lo.y;
    if (tn < lo.z) tn = lo.z;
    if (tn < T(0)) tn = T(0);
    T tf = hi.x;  // std::min({...}): the first smallest
    if (hi.y < tf) tf = hi.y;
    if (hi.z < tf) tf = hi.z;
    if (maxd < tf) tf = maxd;
    if (tf < tn) return false;
    t = tn;
    return true;
}
// ray_vs_triangle (geometry/ray.h:205-226): Moller-Trumbore
template <class T> bool ray_triangle(const V3<T>& o, const V3<T>& dir, T maxd, const V3<T>& a, const V3<T>& b,
                                     const V3<T>& c, T& t) {
    const V3<T> e1 = b - a, e2 = c - a;
    const V3<T> pv = cross(dir, e2);
    const T det = dot(e1, pv);
    if (det > -k_epsilon<T>() && det < k_epsilon<T>()) return false;
    const T inv = T(1) / det;
    const V3<T> tv = o - a;
    const T u = dot(tv, pv) * inv;
    if (u < T(0) || u > T(1)) return false;
    const V3<T> qv = cross(tv, e1);
    const T v = dot(dir, qv) * inv;
    if (v < T(0) || u + v > T(1)) return false;
    const T cand = dot(e2, qv) * inv;
    if (cand < T(0) || cand > maxd) return false;
    t = cand;
    return true;
}
// MeshBVH::raycast (mesh_bvh.h:135-145): leaves of every node the ray's slab test passes, left
// subtree first
template <class T, class Fn> void mesh_bvh_raycast(const Mesh& m, const V3<T>& o, const V3<T>& dir, T maxd, Fn&& fn) {
    if (m.nodes.empty()) return;
    std::vector<int32_t> stack{0};
    while (!stack.empty()) {
        const MeshNode& nd = m.nodes[size_t(stack.back())];
        stack.pop_back();
        AABB<T> b;
        b.mn = mv<T>(nd.b.mn);
        b.mx = mv<T>(nd.b.mx);
        T t;
        if (!ray_box(o, dir, maxd, b, t)) continue;
        if (nd.left < 0) {
            for (uint32_t i = 0; i < nd.count; ++i) fn(m.order[nd.start + i]);
        } else {
            stack.push_back(nd.right);
            stack.push_back(nd.left);
        }
    }
}

// project_at_pose_ (cloth_collision.h:115-360) at the collider's pose with no sweep (the world
// never sets one): moves p to `thickness` outside the shape; normal and depth of the correction.
template <class T> bool project_particle(const Collider<T>& col, T thickness, V3<T>& p, V3<T>& normal, T& depth,
                                         const V3<T>* motion_start) {
    const Shape<T>& s = *col.shape;
    if (!s.enabled) return false;
    const V3<T> pp = col.pos;
    const Q<T> pr = col.rot;
    if (s.type == PXG_SPHERE) {
        const Sph<T> sp = world_sphere(s, pp, pr);
        const V3<T> on_axis = sp.c;
        const V3<T> d = p - on_axis;
        const T len = length(d);
        const T target = sp.r + thickness;
        if (len >= target) return false;
        const V3<T> n = len > k_epsilon<T>() ? d / len : V3<T>(T(0), T(0), T(1));
        depth = target - len;
        p = on_axis + n * target;
        normal = n;
        return true;
    }
    if (s.type == PXG_CAPSULE) {
        const Cap<T> cap = world_capsule(s, pp, pr);
        const V3<T> on_axis = closest_point_on_segment(p, cap.a, cap.b);
        const V3<T> d = p - on_axis;
        const T len = length(d);
        const T target = cap.r + thickness;
        if (len >= target) return false;
        V3<T> n = d;
        if (len > k_epsilon<T>()) {
            n = d / len;
        } else {
            V3<T> t1, t2;
            orthonormal_basis(safe_normalize(cap.b - cap.a), t1, t2);
            n = t1;
        }
        depth = target - len;
        p = on_axis + n * target;
        normal = n;
        return true;
    }
    if (s.type == PXG_BOX) {
        const OBB<T> box = world_obb(s, pp, pr);
        V3<T> ax[3];
        obb_axes(box, ax);
        const V3<T> d = p - box.c;
        const V3<T> local(dot(d, ax[0]), dot(d, ax[1]), dot(d, ax[2]));
        const V3<T> he = box.he;
        const bool inside = std::abs(local.x) <= he.x && std::abs(local.y) <= he.y && std::abs(local.z) <= he.z;
        V3<T> lt, ln;
        if (inside) {
            const V3<T> slack(he.x - std::abs(local.x), he.y - std::abs(local.y), he.z - std::abs(local.z));
            const int axis = (slack.x <= slack.y && slack.x <= slack.z) ? 0 : (slack.y <= slack.z ? 1 : 2);
            const T sign = local[axis] >= T(0) ? T(1) : T(-1);
            lt = local;
            lt[axis] = sign * (he[axis] + thickness);
            ln = V3<T>(T(0));
            ln[axis] = sign;
            depth = slack[axis] + thickness;
        } else {
            lt = V3<T>(gclamp(local.x, -he.x, he.x), gclamp(local.y, -he.y, he.y), gclamp(local.z, -he.z, he.z));
            const V3<T> delta = local - lt;
            const T len = length(delta);
            if (len >= thickness) return false;
            ln = len > k_epsilon<T>() ? delta / len : V3<T>(T(0), T(0), T(1));
            depth = thickness - len;
            lt = lt + ln * thickness;
        }
        p = box.c + ax[0] * lt.x + ax[1] * lt.y + ax[2] * lt.z;
        normal = normalize(ax[0] * ln.x + ax[1] * ln.y + ax[2] * ln.z);
        return true;
    }
    if (s.type == PXG_TRIANGLE_MESH) {
        const Mesh* mesh = static_cast<const Mesh*>(s.mesh);
        if (!mesh || mesh->nodes.empty()) return false;
        const Q<T> rot = qmul(pr, s.lr);
        const Q<T> inv = qinverse(rot);
        const V3<T> origin = pp + qrot(pr, s.lc);
        const T scale = s.mesh_scale > k_epsilon<T>() ? s.mesh_scale : T(1);
        const V3<T> lp = qrot(inv, p - origin) / scale;
        const T lt = thickness / scale;
        auto tri = [&](uint32_t t, V3<T>& a, V3<T>& b, V3<T>& c) {
            a = mv<T>(mesh->v[mesh->idx[3 * t]]);
            b = mv<T>(mesh->v[mesh->idx[3 * t + 1]]);
            c = mv<T>(mesh->v[mesh->idx[3 * t + 2]]);
        };
        if (motion_start) {  // swept crossing test (cloth_collision.h:243-284)
            const V3<T> ls = qrot(inv, *motion_start - origin) / scale;
            const V3<T> seg = lp - ls;
            const T seg_len = length(seg);
            if (seg_len > lt) {
                const V3<T> dir = seg / seg_len;
                T best_t = seg_len;
                uint32_t crossed_tri = 0;
                bool crossed = false;
                mesh_bvh_raycast(*mesh, ls, dir, seg_len, [&](uint32_t t) {
                    V3<T> a, b, c;
                    tri(t, a, b, c);
                    T th = T(0);
                    if ((ray_triangle(ls, dir, seg_len, a, b, c, th) || ray_triangle(ls, dir, seg_len, a, c, b, th)) &&
                        th < best_t) {
                        best_t = th;
                        crossed_tri = t;
                        crossed = true;
                    }
                });
                if (crossed) {
                    const V3<T> hit = ls + dir * best_t;
                    V3<T> face = safe_normalize(mv<T>(mesh->n[crossed_tri]));
                    if (dot(ls - hit, face) < T(0)) face = -face;
                    const V3<T> fixed = hit + face * lt;
                    depth = length(lp - fixed) * scale;
                    p = origin + qrot(rot, fixed * scale);
                    normal = qrot(rot, face);
                    return true;
                }
            }
        }
        AABB<T> q;  // proximity (cloth_collision.h:287-348)
        q.mn = lp - V3<T>(lt);
        q.mx = lp + V3<T>(lt);
        T best_d2 = lt * lt;
        V3<T> best_point(T(0)), best_delta(T(0));
        uint32_t best_tri = 0;
        bool hit = false, any_front = false;
        mesh_bvh_query(*mesh, q, [&](uint32_t t) {
            V3<T> a, b, c;
            tri(t, a, b, c);
            const V3<T> cl = closest_on_triangle(lp, a, b, c);
            const V3<T> delta = lp - cl;
            const T d2 = dot(delta, delta);
            if (d2 > lt * lt) return;
            if (dot(delta, mv<T>(mesh->n[t])) >= T(0)) any_front = true;
            if (d2 > best_d2) return;
            best_d2 = d2;
            best_point = cl;
            best_delta = delta;
            best_tri = t;
            hit = true;
        });
        if (!hit) return false;
        V3<T> bn;
        if (!any_front) {
            bn = safe_normalize(mv<T>(mesh->n[best_tri]));
        } else if (best_d2 > k_epsilon<T>() * k_epsilon<T>()) {
            bn = best_delta / std::sqrt(best_d2);
        } else {
            const V3<T> face = safe_normalize(mv<T>(mesh->n[best_tri]));
            bn = dot(best_delta, face) < T(0) ? -face : face;
        }
        const V3<T> fixed = best_point + bn * lt;
        depth = length(lp - fixed) * scale;
        p = origin + qrot(rot, fixed * scale);
        normal = qrot(rot, bn);
        return true;
    }
    if (s.type == PXG_CONVEX_HULL) {
        if (!s.hull) return false;
        const PosedHull<T> ph = world_posed_hull(s, pp, pr);
        V3<T> point;
        T pen = T(0);
        if (!hull_vs_sphere(ph, Sph<T>{p, thickness}, normal, pen, point)) return false;
        p = point + normal * thickness;
        depth = pen;
        return true;
    }
    return false;
}

// ---- the cloth substep loop (cloth/cloth_solver.h:126-411) ----

template <class T> inline void project_distance(std::vector<CParticle<T>>& ps, CConstraint<T>& k, T alpha) {  // :103-121
    CParticle<T>& pa = ps[k.a];
    CParticle<T>& pb = ps[k.b];
    const T w = pa.inv_mass + pb.inv_mass;
    if (w <= T(0)) return;
    const V3<T> d = pa.pos - pb.pos;
    const T len = length(d);
    if (len < k_epsilon<T>()) return;
    const V3<T> n = d / len;
    const T c = len - k.rest;
    const T dl = (-c - alpha * k.lambda) / (w + alpha);
    k.lambda += dl;
    const V3<T> corr = dl * n;
    pa.pos = pa.pos + pa.inv_mass * corr;
    pb.pos = pb.pos - pb.inv_mass * corr;
}
inline uint32_t cloth_hash_cell(int32_t x, int32_t y, int32_t z, uint32_t buckets) {  // :88-93
    const uint32_t h = uint32_t(x) * 73856093u ^ uint32_t(y) * 19349663u ^ uint32_t(z) * 83492791u;
    return h & (buckets - 1u);
}
inline uint32_t cloth_bucket_count(size_t n) {  // :96-100
    uint32_t b = 16;
    while (b < n * 2u && b < (1u << 22)) b <<= 1;
    return b;
}

template <class T> void solve_cloth(Cloth<T>& c, const std::vector<Collider<T>>& colliders, const V3<T>& gravity,
                                    uint32_t substeps, uint32_t iterations, T h, Atan2Mode mode) {
    const size_t n = c.particles.size();
    if (n == 0 || h <= T(0)) return;
    const uint32_t steps = std::max<uint32_t>(1, c.params.substeps ? c.params.substeps : substeps);
    const uint32_t iters = std::max<uint32_t>(1, c.params.iterations ? c.params.iterations : iterations);
    const T hs = h / T(steps);
    const T inv_hs = T(1) / hs;
    const bool self_col = c.params.self_collision && c.params.self_distance > k_epsilon<T>();
    const uint32_t buckets = self_col ? cloth_bucket_count(n) : 16u;
    const T s_alpha = c.params.stretch_compliance / (hs * hs);
    const T b_alpha = c.params.bend_compliance / (hs * hs);
    V3<T> wind = c.params.wind;
    if (c.params.wind_turbulence > T(0)) {
        const T t = c.time;
        wind = wind + c.params.wind_turbulence *
                          V3<T>(t_sin(t * T(1.7f), mode) * t_cos(t * T(0.9f), mode),
                                t_sin(t * T(1.1f) + T(1.3f), mode) * t_cos(t * T(1.9f), mode),
                                t_sin(t * T(0.7f) + T(2.7f), mode) * t_cos(t * T(1.3f), mode));
    }
    const V3<T> base = gravity * c.params.gravity_scale + c.params.external_acceleration;
    const bool aero = (c.params.air_drag > T(0) || c.params.air_lift > T(0)) && !c.triangles.empty();
    const T maxv2 = c.params.max_velocity * c.params.max_velocity;
    std::vector<V3<T>> accel(n), cnormal(n);
    std::vector<T> cdepth(n);
    std::vector<uint32_t> cell_of(n), counts(buckets + 1), sorted(n);

    for (uint32_t step = 0; step < steps; ++step) {
        std::fill(accel.begin(), accel.end(), V3<T>(T(0)));
        if (aero) {  // 1. aerodynamics
            for (size_t t = 0; t + 2 < c.triangles.size(); t += 3) {
                const uint32_t i0 = c.triangles[t], i1 = c.triangles[t + 1], i2 = c.triangles[t + 2];
                const V3<T>& p0 = c.particles[i0].pos;
                const V3<T> e1 = c.particles[i1].pos - p0, e2 = c.particles[i2].pos - p0;
                const V3<T> cr = cross(e1, e2);
                const T twice = length(cr);
                if (twice < k_epsilon<T>()) continue;
                const V3<T> fn = cr / twice;
                const V3<T> v = (c.particles[i0].vel + c.particles[i1].vel + c.particles[i2].vel) * T(1.0f / 3.0f) - wind;
                const T vn = dot(v, fn);
                const V3<T> vt = v - vn * fn;
                const V3<T> force = T(-0.5f) * twice * std::abs(vn) * (c.params.air_drag * vn * fn + c.params.air_lift * vt);
                const V3<T> share = force * T(1.0f / 3.0f);
                accel[i0] += share;
                accel[i1] += share;
                accel[i2] += share;
            }
        }
        const T damp = T(1) / (T(1) + c.params.damping * hs);  // 2. integrate
        for (size_t i = 0; i < n; ++i) {
            CParticle<T>& p = c.particles[i];
            p.prev = p.pos;
            if (p.inv_mass == T(0)) continue;
            p.vel = p.vel + (base + accel[i] * p.inv_mass) * hs;
            p.vel = p.vel * damp;
            p.pos = p.pos + p.vel * hs;
        }
        const T at = T(step + 1) / T(steps);  // 3. anchors
        for (const CAnchor<T>& a : c.anchors) {
            if (a.particle >= n) continue;
            c.particles[a.particle].pos = vmix(a.prev_world, a.world, at);
        }
        for (auto& k : c.stretch) k.lambda = T(0);  // 4. constraints
        for (auto& k : c.bend) k.lambda = T(0);
        for (uint32_t it = 0; it < iters; ++it) {
            for (const CBatch& b : c.stretch_batches)
                for (uint32_t i = b.begin; i < b.end; ++i) project_distance(c.particles, c.stretch[i], s_alpha);
            for (const CBatch& b : c.bend_batches)
                for (uint32_t i = b.begin; i < b.end; ++i) project_distance(c.particles, c.bend[i], b_alpha);
        }
        for (const CTether<T>& t : c.tethers) {  // 5. tethers
            CParticle<T>& p = c.particles[t.particle];
            if (p.inv_mass == T(0)) continue;
            const V3<T> anchor = c.particles[t.anchor].pos;
            const V3<T> d = p.pos - anchor;
            const T len = length(d);
            if (len <= t.max_len || len < k_epsilon<T>()) continue;
            p.pos = anchor + d * (t.max_len / len);
        }
        if (self_col) {  // 6. self-collision, serial
            const T dmin = c.params.self_distance;
            const T inv_cell = T(1) / dmin;
            std::fill(counts.begin(), counts.end(), 0u);
            for (size_t i = 0; i < n; ++i) {
                const V3<T>& p = c.particles[i].pos;
                const uint32_t cell = cloth_hash_cell(int32_t(std::floor(p.x * inv_cell)), int32_t(std::floor(p.y * inv_cell)),
                                                      int32_t(std::floor(p.z * inv_cell)), buckets);
                cell_of[i] = cell;
                ++counts[cell];
            }
            uint32_t run = 0;
            for (uint32_t b = 0; b < buckets; ++b) {
                const uint32_t k = counts[b];
                counts[b] = run;
                run += k;
            }
            counts[buckets] = run;
            for (size_t i = 0; i < n; ++i) sorted[counts[cell_of[i]]++] = uint32_t(i);
            for (size_t i = 0; i < n; ++i) {
                CParticle<T>& pi = c.particles[i];
                if (pi.inv_mass == T(0)) continue;
                const V3<T> base_c = pi.pos * inv_cell;
                const int32_t cx = int32_t(std::floor(base_c.x)), cy = int32_t(std::floor(base_c.y)),
                              cz = int32_t(std::floor(base_c.z));
                for (int32_t dz = -1; dz <= 1; ++dz)
                    for (int32_t dy = -1; dy <= 1; ++dy)
                        for (int32_t dx = -1; dx <= 1; ++dx) {
                            const uint32_t b = cloth_hash_cell(cx + dx, cy + dy, cz + dz, buckets);
                            const uint32_t first = b == 0 ? 0u : counts[b - 1], last = counts[b];
                            for (uint32_t s = first; s < last; ++s) {
                                const uint32_t j = sorted[s];
                                if (j <= i) continue;
                                if (c.columns > 0) {
                                    const int32_t xi = int32_t(i % c.columns), yi = int32_t(i / c.columns);
                                    const int32_t xj = int32_t(j % c.columns), yj = int32_t(j / c.columns);
                                    if (std::abs(xi - xj) <= 1 && std::abs(yi - yj) <= 1) continue;
                                }
                                CParticle<T>& pj = c.particles[j];
                                const T w = pi.inv_mass + pj.inv_mass;
                                if (w <= T(0)) continue;
                                const V3<T> d = pi.pos - pj.pos;
                                const T l2 = dot(d, d);
                                if (l2 >= dmin * dmin || l2 < k_epsilon<T>()) continue;
                                const T len = std::sqrt(l2);
                                const V3<T> push = (d / len) * ((dmin - len) / w);
                                pi.pos = pi.pos + push * pi.inv_mass;
                                pj.pos = pj.pos - push * pj.inv_mass;
                            }
                        }
            }
        }
        std::fill(cnormal.begin(), cnormal.end(), V3<T>(T(0)));  // 7. collisions
        std::fill(cdepth.begin(), cdepth.end(), T(0));
        if (!colliders.empty()) {
            for (size_t i = 0; i < n; ++i) {
                CParticle<T>& p = c.particles[i];
                if (p.inv_mass == T(0)) continue;
                for (const Collider<T>& col : colliders) {
                    const V3<T>& lo = col.bounds.mn;
                    const V3<T>& hi = col.bounds.mx;
                    const T m = c.params.thickness;
                    if (p.pos.x < lo.x - m || p.pos.x > hi.x + m || p.pos.y < lo.y - m || p.pos.y > hi.y + m ||
                        p.pos.z < lo.z - m || p.pos.z > hi.z + m)
                        continue;
                    V3<T> nrm(T(0));
                    T dep = T(0);
                    if (!project_particle(col, c.params.thickness, p.pos, nrm, dep, &p.prev)) continue;
                    cnormal[i] += nrm;
                    cdepth[i] += dep;
                }
            }
        }
        for (size_t i = 0; i < n; ++i) {  // 8. velocities and friction
            CParticle<T>& p = c.particles[i];
            p.vel = (p.pos - p.prev) * inv_hs;
            const T depth = cdepth[i];
            if (depth > T(0)) {
                const V3<T> ns = cnormal[i];
                const T nl = length(ns);
                if (nl > k_epsilon<T>()) {
                    const V3<T> nrm = ns / nl;
                    const T vn = dot(p.vel, nrm);
                    const V3<T> vt = p.vel - vn * nrm;
                    if (vn < T(0)) p.vel = p.vel - vn * nrm;
                    const T budget = c.params.friction * depth * inv_hs;
                    const T vt_len = length(vt);
                    if (vt_len > k_epsilon<T>()) p.vel = p.vel - vt * std::min(T(1), budget / vt_len);
                }
            }
            const T v2 = dot(p.vel, p.vel);
            if (v2 > maxv2) p.vel = p.vel * (c.params.max_velocity / std::sqrt(v2));
        }
        c.time += hs;
    }
    for (CAnchor<T>& a : c.anchors) a.prev_world = a.world;
    c.bounds = particle_bounds(c.particles);
}

// ---- hair (hair/strand.h, strand_builder.h, strand_solver.h) ----

template <class T> struct Strand {
    uint32_t first = 0, count = 0;
    V3<T> root, prev_root;
    Q<T> root_q, prev_root_q;
};
template <class T> struct StrandParams {
    T stretch_compliance = T(0), global_shape = T(0.22f), global_shape_falloff = T(1.6f), local_shape = T(0.35f);
    T damping = T(0.08f), thickness = T(0.01f), friction = T(0.4f), gravity_scale = T(1), max_velocity = T(50);
    V3<T> external_acceleration, wind;
    T wind_turbulence = T(0.1f), air_drag = T(0.6f);
    uint32_t substeps = 0, iterations = 0;
    T sleep_threshold = T(0.02f), sleep_time = T(0.6f);
    uint32_t layer = 0;
};
template <class T> struct StrandSet {
    std::vector<CParticle<T>> particles;
    std::vector<CConstraint<T>> stretch;
    std::vector<CBatch> stretch_batches;
    std::vector<V3<T>> rest_local, rest_offset;
    std::vector<Strand<T>> strands;
    std::vector<std::pair<uint32_t, uint32_t>> collide_bodies;  // (index, generation)
    StrandParams<T> params;
    AABB<T> bounds;
    bool enabled = true, awake = true;
    T sleep_timer = T(0), time = T(0);
    bool empty() const { return particles.empty() || strands.empty(); }
};
template <class T> struct StrandDesc {
    std::vector<V3<T>> points;
    Q<T> root_q;
    T particle_mass = T(0.001f);
};
// make_strands (strand_builder.h:66-130): root pinned, segments as stretch constraints in two
// particle-disjoint batches (even segments, odd segments), rest pose in the root frame
template <class T> StrandSet<T> make_strands(const std::vector<StrandDesc<T>>& ds, const StrandParams<T>& params) {
    StrandSet<T> set;
    set.params = params;
    size_t total = 0;
    for (const auto& d : ds)
        if (d.points.size() >= 2) total += d.points.size();
    if (total == 0) return set;
    std::vector<CConstraint<T>> even, odd;
    for (const auto& d : ds) {
        if (d.points.size() < 2) continue;
        Strand<T> s;
        s.first = uint32_t(set.particles.size());
        s.count = uint32_t(d.points.size());
        s.root = s.prev_root = d.points.front();
        s.root_q = s.prev_root_q = d.root_q;
        const Q<T> to_root = qinverse(d.root_q);
        const T im = d.particle_mass > T(0) ? T(1) / d.particle_mass : T(1);
        for (size_t k = 0; k < d.points.size(); ++k) {
            CParticle<T> p;
            p.pos = p.prev = d.points[k];
            p.inv_mass = k == 0 ? T(0) : im;
            set.particles.push_back(p);
            set.rest_local.push_back(qrot(to_root, d.points[k] - s.root));
            set.rest_offset.push_back(k == 0 ? V3<T>(T(0)) : d.points[k] - d.points[k - 1]);
            if (k > 0) {
                CConstraint<T> c;
                c.a = s.first + uint32_t(k - 1);
                c.b = s.first + uint32_t(k);
                c.rest = length(d.points[k] - d.points[k - 1]);
                ((k - 1) % 2 == 0 ? even : odd).push_back(c);
            }
        }
        set.strands.push_back(s);
    }
    set.stretch.insert(set.stretch.end(), even.begin(), even.end());
    set.stretch.insert(set.stretch.end(), odd.begin(), odd.end());
    if (!even.empty()) set.stretch_batches.push_back(CBatch{0u, uint32_t(even.size())});
    if (!odd.empty()) set.stretch_batches.push_back(CBatch{uint32_t(even.size()), uint32_t(even.size() + odd.size())});
    set.bounds = particle_bounds(set.particles);
    return set;
}
// set_strand_roots (strand_builder.h:143-153)
template <class T> void set_strand_roots(StrandSet<T>& set, const std::vector<V3<T>>& pos, const std::vector<Q<T>>& rot) {
    const size_t n = std::min(set.strands.size(), pos.size());
    for (size_t i = 0; i < n; ++i) {
        Strand<T>& s = set.strands[i];
        s.prev_root = s.root;
        s.prev_root_q = s.root_q;
        s.root = pos[i];
        if (i < rot.size()) s.root_q = rot[i];
    }
}
// glm::angleAxis(angle, axis): (cos(angle / 2), axis * sin(angle / 2))
template <class T> Q<T> angle_axis(T angle, const V3<T>& axis, Atan2Mode m) {
    const T s = t_sin(angle * T(0.5), m);
    return Q<T>::wxyz(t_cos(angle * T(0.5), m), axis.x * s, axis.y * s, axis.z * s);
}
// glm::slerp (gtc/quaternion.inl, 0.9.9): the short way round, linear within epsilon of 1
template <class T> Q<T> slerp(const Q<T>& x, const Q<T>& y, T a, Atan2Mode m) {
    Q<T> z = y;
    T ct = qdot(x, y);
    if (ct < T(0)) {
        z = Q<T>::wxyz(-y.w, -y.x, -y.y, -y.z);
        ct = -ct;
    }
    if (ct > T(1) - std::numeric_limits<T>::epsilon())
        return Q<T>::wxyz(smix(x.w, z.w, a), smix(x.x, z.x, a), smix(x.y, z.y, a), smix(x.z, z.z, a));
    const T angle = t_acos(ct, m);
    const Q<T> sum = qadd(qscale(x, t_sin((T(1) - a) * angle, m)), qscale(z, t_sin(a * angle, m)));
    const T d = t_sin(angle, m);
    return Q<T>::wxyz(sum.w / d, sum.x / d, sum.y / d, sum.z / d);
}
// detail::rotation_between (strand_solver.h:70-85)
template <class T> Q<T> rotation_between(const V3<T>& from, const V3<T>& to, Atan2Mode m) {
    const T d = gclamp(dot(from, to), T(-1), T(1));
    if (d > T(0.999999f)) return Q<T>();
    if (d < T(-0.999999f)) {
        V3<T> axis = cross(V3<T>(T(0), T(0), T(1)), from);
        if (dot(axis, axis) < T(1e-8f)) axis = cross(V3<T>(T(1), T(0), T(0)), from);
        return angle_axis(T(3.14159265358979323846f), normalize(axis), m);
    }
    const V3<T> axis = cross(from, to);
    const T len = length(axis);
    if (len < T(1e-8f)) return Q<T>();
    return angle_axis(t_acos(d, m), axis / len, m);
}

// solve_strands (strand_solver.h:121-317)
template <class T> void solve_strands(StrandSet<T>& set, const std::vector<Collider<T>>& colliders, const V3<T>& gravity,
                                      uint32_t substeps, uint32_t iterations, T h, Atan2Mode mode) {
    const size_t n = set.particles.size();
    if (n == 0 || set.strands.empty() || h <= T(0)) return;
    const uint32_t steps = std::max<uint32_t>(1, set.params.substeps ? set.params.substeps : substeps);
    const uint32_t iters = std::max<uint32_t>(1, set.params.iterations ? set.params.iterations : iterations);
    const T hs = h / T(steps);
    const T inv_hs = T(1) / hs;
    const T alpha = set.params.stretch_compliance * inv_hs * inv_hs;
    const V3<T> accel = gravity * set.params.gravity_scale + set.params.external_acceleration;
    const T maxv2 = set.params.max_velocity * set.params.max_velocity;
    const T frac = hs / T(1.0f / 60.0f);
    const T g_sub = T(1) - t_pow(T(1) - gclamp(set.params.global_shape, T(0), T(1)), frac, mode);
    const T l_sub = T(1) - t_pow(T(1) - gclamp(set.params.local_shape, T(0), T(1)), frac, mode);
    std::vector<V3<T>> cnormal(n);
    std::vector<T> cdepth(n);
    std::vector<uint8_t> contacted(n);
    for (uint32_t s = 0; s < steps; ++s) {
        const T t = T(s + 1) / T(steps);
        for (size_t i = 0; i < n; ++i) {  // 1 + 2: aerodynamics and integration
            CParticle<T>& p = set.particles[i];
            if (p.inv_mass <= T(0)) continue;
            V3<T> a = accel;
            if (set.params.air_drag > T(0)) {
                V3<T> wind = set.params.wind;
                if (set.params.wind_turbulence > T(0)) {
                    const T phase = set.time * T(2.3f) + p.pos.x * T(1.7f) + p.pos.y * T(2.1f) + p.pos.z * T(1.3f);
                    wind = wind + set.params.wind_turbulence * length(set.params.wind) *
                                      V3<T>(t_sin(phase, mode), t_sin(phase * T(1.37f) + T(2.1f), mode),
                                            t_sin(phase * T(0.81f) + T(4.2f), mode));
                }
                a = a + (wind - p.vel) * set.params.air_drag;
            }
            p.vel = p.vel + a * hs;
            p.vel = p.vel * (T(1) / (T(1) + set.params.damping * hs));
            const T v2 = dot(p.vel, p.vel);
            if (v2 > maxv2) p.vel = p.vel * (set.params.max_velocity / std::sqrt(v2));
            p.prev = p.pos;
            p.pos = p.pos + p.vel * hs;
        }
        for (Strand<T>& st : set.strands) {  // 3: roots
            if (st.count == 0) continue;
            CParticle<T>& p = set.particles[st.first];
            p.prev = p.pos;
            p.pos = vmix(st.prev_root, st.root, t);
            p.vel = V3<T>(T(0));
        }
        for (auto& k : set.stretch) k.lambda = T(0);  // 4: stretch
        for (uint32_t it = 0; it < iters; ++it)
            for (const CBatch& b : set.stretch_batches)
                for (uint32_t ci = b.begin; ci < b.end; ++ci) project_distance(set.particles, set.stretch[ci], alpha);
        if (set.params.global_shape > T(0)) {  // 5: global shape
            for (const Strand<T>& st : set.strands) {
                if (st.count < 2) continue;
                const Q<T> q = slerp(st.prev_root_q, st.root_q, t, mode);
                const V3<T> root = vmix(st.prev_root, st.root, t);
                const T inv_last = T(1) / T(st.count - 1);
                for (uint32_t k = 1; k < st.count; ++k) {
                    const size_t i = st.first + k;
                    CParticle<T>& p = set.particles[i];
                    if (p.inv_mass <= T(0)) continue;
                    const T along = T(k) * inv_last;
                    const T weight = g_sub * t_pow(T(1) - along, set.params.global_shape_falloff, mode);
                    if (weight <= T(0)) continue;
                    const V3<T> target = root + qrot(q, set.rest_local[i]);
                    p.pos = vmix(p.pos, target, gmin(weight, T(1)));
                }
            }
        }
        if (set.params.local_shape > T(0)) {  // 6: local shape
            for (const Strand<T>& st : set.strands) {
                if (st.count < 3) continue;
                for (uint32_t k = 1; k + 1 < st.count; ++k) {
                    const size_t prev = st.first + k - 1, cur = st.first + k, next = st.first + k + 1;
                    const V3<T> rest_prev = set.rest_offset[cur];
                    const V3<T> live_prev = set.particles[cur].pos - set.particles[prev].pos;
                    const T rl = length(rest_prev), ll = length(live_prev);
                    if (rl <= T(1e-8f) || ll <= T(1e-8f)) continue;
                    const Q<T> turn = rotation_between(rest_prev / rl, live_prev / ll, mode);
                    const V3<T> target = set.particles[cur].pos + qrot(turn, set.rest_offset[next]);
                    const V3<T> delta = (target - set.particles[next].pos) * l_sub;
                    CParticle<T>& a = set.particles[cur];
                    CParticle<T>& b = set.particles[next];
                    const T w = a.inv_mass + b.inv_mass;
                    if (w <= T(0)) continue;
                    b.pos = b.pos + delta * (b.inv_mass / w);
                    a.pos = a.pos - delta * (a.inv_mass / w);
                }
            }
        }
        std::fill(contacted.begin(), contacted.end(), uint8_t(0));  // 7: collisions
        if (!colliders.empty()) {
            for (size_t i = 0; i < n; ++i) {
                CParticle<T>& p = set.particles[i];
                if (p.inv_mass <= T(0)) continue;
                for (const Collider<T>& col : colliders) {
                    V3<T> nrm(T(0));
                    T dep = T(0);
                    if (project_particle(col, set.params.thickness, p.pos, nrm, dep, &p.prev)) {
                        if (contacted[i] == 0 || dep > cdepth[i]) {
                            cnormal[i] = nrm;
                            cdepth[i] = dep;
                        }
                        contacted[i] = 1;
                    }
                }
            }
        }
        for (size_t i = 0; i < n; ++i) {  // 8: velocities and friction
            CParticle<T>& p = set.particles[i];
            if (p.inv_mass <= T(0)) continue;
            p.vel = (p.pos - p.prev) * inv_hs;
            if (contacted[i] != 0 && set.params.friction > T(0)) {
                const V3<T>& nrm = cnormal[i];
                const V3<T> motion = p.pos - p.prev;
                const V3<T> tang = motion - nrm * dot(motion, nrm);
                const T tl = length(tang);
                if (tl > k_epsilon<T>()) {
                    const T limit = set.params.friction * cdepth[i];
                    const T scale = gmax(T(0), T(1) - gmin(limit, tl) / tl);
                    p.pos = p.pos - tang * (T(1) - scale);
                    p.vel = (p.pos - p.prev) * inv_hs;
                }
            }
        }
        set.time += hs;
    }
    set.bounds = particle_bounds(set.particles);
}

}  // namespace pxo
