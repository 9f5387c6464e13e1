// ORACLE -- test infrastructure only. Nothing in the product may include, link or run this.
//
// extern "C" surface of the CPU oracle, loaded by tests/ (ctypes) and by bench.py's CPU
// baseline leg. Mirrors the pxg_* C ABI so a test can drive both sides with the same buffers.
#include "pxo_step.hpp"
#include "pxo_query.hpp"

#include <chrono>
#include <map>

using namespace pxo;

namespace {

template <class T> void cfg_from(Config<T>& c, const pxg_config* in) {
    c.vel_iters = in->velocity_iterations; c.relax_iters = in->relax_iterations; c.pos_iters = in->position_iterations;
    c.pos_corr = T(in->position_correction); c.slop = T(in->linear_slop); c.max_corr = T(in->max_linear_correction);
    c.rest_thresh = T(in->restitution_threshold); c.sleep_lin = T(in->sleep_linear); c.sleep_ang = T(in->sleep_angular);
    c.sleep_time = T(in->sleep_time); c.margin = T(in->aabb_margin); c.max_substeps = in->max_substeps;
    c.fixed_dt = T(in->fixed_dt); c.max_lin_vel = T(in->max_linear_velocity); c.exact_quat = in->use_exact_quaternion_integration != 0;
    c.gravity = V3<T>(T(in->gravity[0]), T(in->gravity[1]), T(in->gravity[2]));
    c.residual_mode = in->residual_mode != 0; c.residual_tol = T(in->residual_tolerance); c.residual_max = in->residual_max_iterations;
    c.residual_every = in->residual_check_interval ? in->residual_check_interval : 1u;
    c.cloth_substeps = in->cloth_substeps; c.cloth_iterations = in->cloth_iterations;
}
template <class T> V3<T> v3(const float* f) { return V3<T>(T(f[0]), T(f[1]), T(f[2])); }
template <class T> Q<T> q4(const float* f) { Q<T> q; q.x = T(f[0]); q.y = T(f[1]); q.z = T(f[2]); q.w = T(f[3]); return q; }
template <class T> void put3(float* f, const V3<T>& v) { f[0] = float(v.x); f[1] = float(v.y); f[2] = float(v.z); }
template <class T> void put4(float* f, const Q<T>& q) { f[0] = float(q.x); f[1] = float(q.y); f[2] = float(q.z); f[3] = float(q.w); }

template <class T> Body<T> body_from(const pxg_body& d) {
    Body<T> b;
    b.type = int(d.type); b.alive = d.alive != 0; b.awake = d.awake != 0; b.use_gravity = d.use_gravity != 0;
    b.constraints = d.constraints; b.generation = d.generation;
    b.pos = v3<T>(d.position); b.sleep_timer = T(d.sleep_timer); b.q = q4<T>(d.orientation);
    b.v = v3<T>(d.linear_velocity); b.mass = T(d.mass); b.w = v3<T>(d.angular_velocity); b.inv_mass = T(d.inv_mass);
    b.force = v3<T>(d.force); b.ldrag = T(d.linear_drag); b.torque = v3<T>(d.torque); b.adrag = T(d.angular_drag);
    b.iil = v3<T>(d.inv_inertia_local);
    b.prev_pos = v3<T>(d.prev_position); b.prev_q = q4<T>(d.prev_orientation);
    update_world_inertia(b);  // add_body(), world.h:113
    return b;
}
template <class T> void body_to(const Body<T>& b, pxg_body& d) {
    std::memset(&d, 0, sizeof(d));
    d.type = uint32_t(b.type); d.alive = b.alive; d.awake = b.awake; d.use_gravity = b.use_gravity;
    d.constraints = b.constraints; d.generation = b.generation;
    put3(d.position, b.pos); d.sleep_timer = float(b.sleep_timer); put4(d.orientation, b.q);
    put3(d.linear_velocity, b.v); d.mass = float(b.mass); put3(d.angular_velocity, b.w); d.inv_mass = float(b.inv_mass);
    put3(d.force, b.force); d.linear_drag = float(b.ldrag); put3(d.torque, b.torque); d.angular_drag = float(b.adrag);
    put3(d.inv_inertia_local, b.iil);
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) d.inv_inertia_world[c * 3 + r] = float(b.iiw.c[c][r]);
    put3(d.prev_position, b.prev_pos); put4(d.prev_orientation, b.prev_q);
}
template <class T> Shape<T> shape_from(const pxg_shape& d) {
    Shape<T> s;
    s.owner = d.owner; s.type = int(d.type); s.enabled = d.enabled != 0; s.trigger = d.is_trigger != 0;
    s.layer = d.layer; s.caxis = d.capsule_axis; s.fcomb = int(d.friction_combine); s.rcomb = int(d.restitution_combine);
    s.lc = v3<T>(d.local_center); s.radius = T(d.radius); s.lr = q4<T>(d.local_rotation); s.he = v3<T>(d.half_extents);
    s.cr = T(d.capsule_radius); s.chh = T(d.capsule_half_height); s.friction = T(d.dynamic_friction); s.restitution = T(d.restitution);
    if (s.type == PXG_CONVEX_HULL) s.hull_scale = T(d.radius);
    if (s.type == PXG_TRIANGLE_MESH) s.mesh_scale = T(d.radius);
    return s;
}
// a mesh shape's pointer into the world's mesh table (nullptr for a bad index: no contacts)
template <class T> void resolve_mesh(Shape<T>& s, const std::vector<Mesh>& meshes, int32_t index) {
    if (s.type != PXG_TRIANGLE_MESH) return;
    s.mesh = (index >= 0 && size_t(index) < meshes.size()) ? &meshes[size_t(index)] : nullptr;
}
// from_triangles of a pxg_mesh_desc (false for an index out of range)
inline bool mesh_from(const pxg_mesh_desc& d, Mesh& out) {
    std::vector<V3<float>> pos(d.n_positions);
    for (uint32_t i = 0; i < d.n_positions; ++i) pos[i] = V3<float>(d.positions[3 * i], d.positions[3 * i + 1], d.positions[3 * i + 2]);
    std::vector<uint32_t> ind(d.indices, d.indices + d.n_indices);
    for (uint32_t x : ind)
        if (x >= d.n_positions) return false;
    out = mesh_from_triangles(pos, ind);
    return true;
}
// a hull shape's pointer into the world's hull table (nullptr for a bad index: no contacts)
template <class T> void resolve_hull(Shape<T>& s, const std::vector<Hull<T>>& hulls, int32_t index) {
    if (s.type != PXG_CONVEX_HULL) return;
    s.hull = (index >= 0 && size_t(index) < hulls.size()) ? &hulls[size_t(index)] : nullptr;
}
inline Hull<float> hull_from(const pxg_hull& d) {
    Hull<float> h;
    for (uint32_t i = 0; i < d.n_vertices; ++i) h.v.push_back(V3<float>(d.vertices[3 * i], d.vertices[3 * i + 1], d.vertices[3 * i + 2]));
    for (uint32_t f = 0; f < d.n_faces; ++f) {
        h.pn.push_back(V3<float>(d.face_planes[4 * f], d.face_planes[4 * f + 1], d.face_planes[4 * f + 2]));
        h.pd.push_back(d.face_planes[4 * f + 3]);
    }
    h.fs.assign(d.face_starts, d.face_starts + d.n_faces + 1);
    h.fv.assign(d.face_vertices, d.face_vertices + d.face_starts[d.n_faces]);
    for (uint32_t e = 0; e < d.n_edges; ++e) h.e.push_back({d.edges[2 * e], d.edges[2 * e + 1]});
    h.lb.mn = V3<float>(d.local_min[0], d.local_min[1], d.local_min[2]);
    h.lb.mx = V3<float>(d.local_max[0], d.local_max[1], d.local_max[2]);
    h.centroid = V3<float>(d.centroid[0], d.centroid[1], d.centroid[2]);
    h.volume = d.volume;
    return h;
}
template <class T> void manifold_to(const Manifold<T>& m, pxg_manifold& d) {
    std::memset(&d, 0, sizeof(d));
    d.a = m.a; d.b = m.b; d.shape_a = m.sa; d.shape_b = m.sb;
    put3(d.normal, m.n); d.friction = float(m.friction); d.restitution = float(m.restitution);
    d.count = uint32_t(m.count); d.is_trigger = m.trigger; d.valid = m.valid;
    for (int i = 0; i < 4; ++i) {
        const Contact<T>& c = m.p[i];
        pxg_contact& o = d.points[i];
        put3(o.position, c.pos); o.penetration = float(c.pen); o.separation = float(c.sep);
        o.normal_impulse = float(c.ni); o.tangent_impulse[0] = float(c.ti[0]); o.tangent_impulse[1] = float(c.ti[1]);
        o.feature_id = c.fid; o.restitution_bias = float(c.rbias);
    }
}

template <class T> Manifold<T> manifold_from(const pxg_manifold& d) {
    Manifold<T> m;
    m.a = d.a; m.b = d.b; m.sa = d.shape_a; m.sb = d.shape_b;
    m.n = v3<T>(d.normal); m.friction = T(d.friction); m.restitution = T(d.restitution);
    m.count = int(d.count); m.trigger = d.is_trigger != 0; m.valid = d.valid != 0;
    for (int i = 0; i < 4; ++i) {
        const pxg_contact& o = d.points[i];
        Contact<T>& c = m.p[i];
        c.pos = v3<T>(o.position); c.pen = T(o.penetration); c.sep = T(o.separation);
        c.ni = T(o.normal_impulse); c.ti[0] = T(o.tangent_impulse[0]); c.ti[1] = T(o.tangent_impulse[1]);
        c.fid = o.feature_id; c.rbias = T(o.restitution_bias);
    }
    return m;
}

template <class T> Joint<T> joint_from(const pxg_joint& d) {
    Joint<T> j;
    j.type = int(d.type); j.a = d.body_a; j.b = d.body_b;
    j.valid = d.valid != 0; j.enabled = d.enabled != 0; j.use_limits = d.use_limits != 0;
    j.collide_connected = d.collide_connected != 0;
    j.anchor_a = v3<T>(d.local_anchor_a); j.anchor_b = v3<T>(d.local_anchor_b);
    j.axis_a = v3<T>(d.local_axis_a); j.axis_b = v3<T>(d.local_axis_b);
    j.min_angle = T(d.min_angle); j.max_angle = T(d.max_angle);
    j.swing_limit = T(d.swing_limit); j.twist_min = T(d.twist_min); j.twist_max = T(d.twist_max);
    j.rest = q4<T>(d.rest_relative_rotation);
    j.point_impulse = v3<T>(d.point_impulse);
    j.axis_impulse[0] = T(d.axis_impulse[0]); j.axis_impulse[1] = T(d.axis_impulse[1]);
    j.limit_impulse = T(d.limit_impulse); j.swing_impulse = T(d.swing_impulse);
    return j;
}
template <class T> void joint_to(const Joint<T>& j, pxg_joint& d) {
    std::memset(&d, 0, sizeof(d));
    d.type = uint32_t(j.type); d.body_a = j.a; d.body_b = j.b;
    d.valid = j.valid; d.enabled = j.enabled; d.use_limits = j.use_limits; d.collide_connected = j.collide_connected;
    put3(d.local_anchor_a, j.anchor_a); put3(d.local_anchor_b, j.anchor_b);
    put3(d.local_axis_a, j.axis_a); put3(d.local_axis_b, j.axis_b);
    d.min_angle = float(j.min_angle); d.max_angle = float(j.max_angle);
    d.swing_limit = float(j.swing_limit); d.twist_min = float(j.twist_min); d.twist_max = float(j.twist_max);
    put4(d.rest_relative_rotation, j.rest);
    put3(d.point_impulse, j.point_impulse);
    d.axis_impulse[0] = float(j.axis_impulse[0]); d.axis_impulse[1] = float(j.axis_impulse[1]);
    d.limit_impulse = float(j.limit_impulse); d.swing_impulse = float(j.swing_impulse);
}

struct Handle {
    int dbl = 0;
    World<float> f;
    World<double> d;
};

template <class T> World<T>& W(Handle* h);
template <> World<float>& W<float>(Handle* h) { return h->f; }
template <> World<double>& W<double>(Handle* h) { return h->d; }

// events(), clear_events() (world.h:744-748); out4: a, b (owner body indices), type
// (0 enter, 1 stay, 2 exit), is_trigger per event
template <class T> uint32_t get_events_t(World<T>& w, uint32_t* out4, uint32_t cap) {
    uint32_t n = uint32_t(w.events.size());
    for (uint32_t i = 0; i < n && i < cap; ++i) {
        out4[4 * i] = w.events[i].a; out4[4 * i + 1] = w.events[i].b;
        out4[4 * i + 2] = w.events[i].type; out4[4 * i + 3] = w.events[i].trigger;
    }
    return n;
}
template <class T> Ray<T> ray_from(const pxg_ray& r) {
    Ray<T> o;
    o.o = v3<T>(r.origin);
    o.d = v3<T>(r.direction);
    o.max_d = T(r.max_distance);
    return o;
}
inline QueryFilter filter_from(const pxg_query_filter* f) {
    QueryFilter q;
    if (f) { q.layer_mask = f->layer_mask; q.include_triggers = f->include_triggers != 0; q.ignore = f->ignore_body; }
    return q;
}
template <class T> OverlapQuery<T> overlap_from(const pxg_overlap_query& in) {
    OverlapQuery<T> q;
    q.type = int(in.type);
    q.sph.c = v3<T>(in.center); q.sph.r = T(in.radius);
    q.box.c = v3<T>(in.center); q.box.he = v3<T>(in.half_extents); q.box.q = q4<T>(in.orientation);
    q.cap.a = v3<T>(in.a); q.cap.b = v3<T>(in.b); q.cap.r = T(in.radius);
    return q;
}
template <class T> void hit_to(const RayHit<T>& h, pxg_raycast_hit& o) {
    std::memset(&o, 0, sizeof(o));
    put3(o.point, h.point); o.distance = float(h.distance); put3(o.normal, h.normal);
    o.hit = 1; o.body = h.body; o.shape_index = h.shape;
}

#define DISPATCH(h, call) ((h)->dbl ? call<double>(W<double>(h)) : call<float>(W<float>(h)))

}  // namespace

extern "C" {

void* pxo_create(const pxg_config* cfg, int dbl) {
    Handle* h = new Handle();
    h->dbl = dbl;
    cfg_from(h->f.cfg, cfg);
    cfg_from(h->d.cfg, cfg);
    return h;
}
void pxo_destroy(void* p) { delete static_cast<Handle*>(p); }

void pxo_set_config(void* p, const pxg_config* cfg) {
    Handle* h = static_cast<Handle*>(p);
    cfg_from(h->f.cfg, cfg);
    cfg_from(h->d.cfg, cfg);
}

void pxo_set_bodies(void* p, const pxg_body* in, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        w.bodies.resize(n);
        for (uint32_t i = 0; i < n; ++i) w.bodies[i] = body_from<T>(in[i]);
    };
    if (h->dbl) go(h->d); else go(h->f);
}
void pxo_get_bodies(void* p, pxg_body* out, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) { for (uint32_t i = 0; i < n && i < w.bodies.size(); ++i) body_to(w.bodies[i], out[i]); };
    if (h->dbl) go(h->d); else go(h->f);
}
void pxo_set_shapes(void* p, const pxg_shape* in, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        w.shapes.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            w.shapes[i] = shape_from<T>(in[i]);
            resolve_hull(w.shapes[i], w.hulls, in[i].capsule_axis);
            resolve_mesh(w.shapes[i], w.meshes, in[i].capsule_axis);
        }
        w.create_proxies();
    };
    if (h->dbl) go(h->d); else go(h->f);
}
void pxo_set_layer_masks(void* p, const uint32_t* m) {
    Handle* h = static_cast<Handle*>(p);
    for (int i = 0; i < 32; ++i) { h->f.layer_mask[i] = m[i]; h->d.layer_mask[i] = m[i]; }
}
void pxo_set_joints(void* p, const pxg_joint* in, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        w.joints.resize(n);
        for (uint32_t i = 0; i < n; ++i) w.joints[i] = joint_from<T>(in[i]);
    };
    if (h->dbl) go(h->d); else go(h->f);
}
uint32_t pxo_get_joints(void* p, pxg_joint* out, uint32_t cap) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) -> uint32_t {
        for (uint32_t i = 0; i < cap && i < w.joints.size(); ++i) joint_to(w.joints[i], out[i]);
        return uint32_t(w.joints.size());
    };
    return h->dbl ? go(h->d) : go(h->f);
}
void pxo_set_ignored_pairs(void* p, const uint32_t* ab, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        w.ignored.clear();
        for (uint32_t i = 0; i < n; ++i) w.ignored.insert(w.body_pair_key(ab[2 * i], ab[2 * i + 1]));
    };
    if (h->dbl) go(h->d); else go(h->f);
}
// 0: std::atan2 on the scalar type (the reference's call); 1: atan2 in double, rounded once
void pxo_set_atan2_mode(void* p, uint32_t mode) {
    Handle* h = static_cast<Handle*>(p);
    h->f.atan2_mode = mode ? Atan2Mode::Rounded : Atan2Mode::Libm;
    h->d.atan2_mode = h->f.atan2_mode;
}
// The hull table PXG_CONVEX_HULL shapes index (upload before the shapes).
void pxo_set_hulls(void* p, const pxg_hull* in, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    h->f.hulls.clear();
    h->d.hulls.clear();
    for (uint32_t i = 0; i < n; ++i) {
        Hull<float> hf = hull_from(in[i]);
        h->f.hulls.push_back(hf);
        h->d.hulls.push_back(hf.as<double>());
    }
}
// The mesh table PXG_TRIANGLE_MESH shapes index (upload before the shapes); 0 for a bad index.
int pxo_set_meshes(void* p, const pxg_mesh_desc* in, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    std::vector<Mesh> ms(n);
    for (uint32_t i = 0; i < n; ++i)
        if (!mesh_from(in[i], ms[i])) return 0;
    h->f.meshes = ms;
    h->d.meshes = ms;
    return 1;
}
// from_triangles into caller arrays (pxg_cook_mesh's layout) plus the tree: nodes as
// (min xyz, max xyz) floats and (left, right, start, count) ints, and the triangle order.
int pxo_cook_mesh(const pxg_mesh_desc* in, pxg_mesh_info* info, float* vertices, uint32_t* indices, float* normals,
                  uint32_t* adjacency, float* node_bounds, int32_t* node_links, uint32_t* order) {
    Mesh m;
    if (!mesh_from(*in, m)) return 0;
    info->n_vertices = uint32_t(m.v.size());
    info->n_triangles = uint32_t(m.tris());
    info->n_nodes = uint32_t(m.nodes.size());
    info->non_manifold_edges = uint32_t(m.non_manifold);
    put3(info->local_min, m.bounds.mn);
    put3(info->local_max, m.bounds.mx);
    for (size_t i = 0; vertices && i < m.v.size(); ++i) put3(vertices + 3 * i, m.v[i]);
    for (size_t i = 0; indices && i < m.idx.size(); ++i) indices[i] = m.idx[i];
    for (size_t t = 0; normals && t < m.tris(); ++t) put3(normals + 3 * t, m.n[t]);
    for (size_t t = 0; adjacency && t < m.tris(); ++t)
        for (int e = 0; e < 3; ++e) adjacency[3 * t + size_t(e)] = m.adj[t][size_t(e)];
    for (size_t k = 0; node_bounds && k < m.nodes.size(); ++k) {
        put3(node_bounds + 6 * k, m.nodes[k].b.mn);
        put3(node_bounds + 6 * k + 3, m.nodes[k].b.mx);
    }
    for (size_t k = 0; node_links && k < m.nodes.size(); ++k) {
        node_links[4 * k] = m.nodes[k].left; node_links[4 * k + 1] = m.nodes[k].right;
        node_links[4 * k + 2] = int32_t(m.nodes[k].start); node_links[4 * k + 3] = int32_t(m.nodes[k].count);
    }
    for (size_t t = 0; order && t < m.order.size(); ++t) order[t] = m.order[t];
    return 1;
}
// build_convex_hull (convex_hull.h:188-449) of n points (xyz). Returns 1 when a hull was built;
// sizes[0..3] = vertices, faces, face vertex entries, edges; the arrays receive at most the caps
// given in sizes[0..3] on entry. volume_centroid: volume, centroid xyz.
int pxo_build_hull(const float* pts, uint32_t n, float tol, uint32_t max_vertices, uint32_t* sizes, float* vertices,
                   float* planes, uint32_t* face_starts, uint32_t* face_vertices, uint32_t* edges, float* bounds6,
                   float* volume_centroid) {
    std::vector<V3<float>> cloud(n);
    for (uint32_t i = 0; i < n; ++i) cloud[i] = V3<float>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    Hull<float> hl;
    bool ok = build_convex_hull(cloud.data(), cloud.size(), hl, tol, max_vertices);
    const uint32_t cap[4] = {sizes[0], sizes[1], sizes[2], sizes[3]};
    sizes[0] = uint32_t(hl.v.size()); sizes[1] = uint32_t(hl.faces()); sizes[2] = uint32_t(hl.fv.size());
    sizes[3] = uint32_t(hl.e.size());
    if (!ok) return 0;
    if (sizes[0] > cap[0] || sizes[1] > cap[1] || sizes[2] > cap[2] || sizes[3] > cap[3]) return 1;
    for (size_t i = 0; i < hl.v.size(); ++i) { vertices[3 * i] = hl.v[i].x; vertices[3 * i + 1] = hl.v[i].y; vertices[3 * i + 2] = hl.v[i].z; }
    for (size_t f = 0; f < hl.faces(); ++f) {
        planes[4 * f] = hl.pn[f].x; planes[4 * f + 1] = hl.pn[f].y; planes[4 * f + 2] = hl.pn[f].z; planes[4 * f + 3] = hl.pd[f];
    }
    for (size_t f = 0; f <= hl.faces(); ++f) face_starts[f] = hl.fs[f];
    for (size_t k = 0; k < hl.fv.size(); ++k) face_vertices[k] = hl.fv[k];
    for (size_t e = 0; e < hl.e.size(); ++e) { edges[2 * e] = hl.e[e].first; edges[2 * e + 1] = hl.e[e].second; }
    bounds6[0] = hl.lb.mn.x; bounds6[1] = hl.lb.mn.y; bounds6[2] = hl.lb.mn.z;
    bounds6[3] = hl.lb.mx.x; bounds6[4] = hl.lb.mx.y; bounds6[5] = hl.lb.mx.z;
    volume_centroid[0] = hl.volume;
    volume_centroid[1] = hl.centroid.x; volume_centroid[2] = hl.centroid.y; volume_centroid[3] = hl.centroid.z;
    return 1;
}
// Host threads for bounds, pair discovery and narrowphase (1: serial, the reference's default).
void pxo_set_threads(void* p, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    if (h->dbl) h->d.set_threads(n);
    else h->f.set_threads(n);
}
void pxo_set_brute_force(void* p, uint32_t on) {
    Handle* h = static_cast<Handle*>(p);
    h->f.brute_force = on; h->d.brute_force = on;
}
// Solve order for the following steps: n shape pairs (a, b).  n == 0 clears it (pair order).
void pxo_set_order(void* p, const uint32_t* ab, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        w.order_hint.clear();
        for (uint32_t i = 0; i < n; ++i) w.order_hint.push_back({ab[2 * i], ab[2 * i + 1]});
        w.use_order_hint = n > 0;
    };
    if (h->dbl) go(h->d); else go(h->f);
}
void pxo_set_miss_table(void* p, const uint64_t* keys, const float* imp, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        w.miss_table.clear();
        for (uint32_t i = 0; i < n; ++i) { Cached c; c.n = imp[3 * i]; c.t1 = imp[3 * i + 1]; c.t2 = imp[3 * i + 2]; w.miss_table[keys[i]] = c; }
        w.use_miss_table = n > 0;
    };
    if (h->dbl) go(h->d); else go(h->f);
}
void pxo_step_fixed(void* p, float hstep) {
    Handle* h = static_cast<Handle*>(p);
    if (h->dbl) h->d.step_fixed(double(hstep)); else h->f.step_fixed(hstep);
}
void pxo_step_phase(void* p, uint32_t phase, float hstep) {
    Handle* h = static_cast<Handle*>(p);
    if (h->dbl) { if (phase == 0) h->d.step_phase_collide(double(hstep)); else h->d.step_phase_solve(double(hstep)); }
    else { if (phase == 0) h->f.step_phase_collide(hstep); else h->f.step_phase_solve(hstep); }
}
void pxo_clear_events(void* p);
// Runs `steps` fixed steps and returns the wall time of the loop in seconds (steady_clock).
double pxo_run(void* p, uint32_t steps, float hstep) {
    Handle* h = static_cast<Handle*>(p);
    auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < steps; ++i) {
        pxo_clear_events(p);  // one step() per frame clears the queue first (world.h:1230)
        pxo_step_fixed(p, hstep);
    }
    auto t1 = std::chrono::steady_clock::now();
    (void)h;
    return std::chrono::duration<double>(t1 - t0).count();
}
// raycast() / raycast_any() (closest = 0: any hit, reported as the closest) for n rays;
// raycast_all() for one ray (returns the hit count)
void pxo_raycast(void* p, const pxg_ray* rays, uint32_t n, const pxg_query_filter* f, pxg_raycast_hit* out) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        for (uint32_t i = 0; i < n; ++i) {
            auto hits = world_raycast(w, ray_from<T>(rays[i]), filter_from(f), true);
            if (hits.empty()) std::memset(&out[i], 0, sizeof(out[i]));
            else hit_to(hits[0], out[i]);
        }
    };
    if (h->dbl) go(h->d); else go(h->f);
}
uint32_t pxo_raycast_all(void* p, const pxg_ray* ray, const pxg_query_filter* f, pxg_raycast_hit* out, uint32_t cap) {
    Handle* h = static_cast<Handle*>(p);
    uint32_t n = 0;
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        auto hits = world_raycast(w, ray_from<T>(*ray), filter_from(f), false);
        n = uint32_t(hits.size());
        for (uint32_t i = 0; i < n && i < cap; ++i) hit_to(hits[i], out[i]);
    };
    if (h->dbl) go(h->d); else go(h->f);
    return n;
}
// overlap_sphere / overlap_box / overlap_capsule: returns the count, copies min(cap, count) bodies
uint32_t pxo_overlap(void* p, const pxg_overlap_query* q, const pxg_query_filter* f, uint32_t* bodies, uint32_t cap) {
    Handle* h = static_cast<Handle*>(p);
    uint32_t n = 0;
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        auto out = world_overlap(w, overlap_from<T>(*q), filter_from(f));
        n = uint32_t(out.size());
        for (uint32_t i = 0; i < n && i < cap; ++i) bodies[i] = out[i];
    };
    if (h->dbl) go(h->d); else go(h->f);
    return n;
}
// shape_cast() for a sphere, capsule or box (returns 1 on a hit)
int pxo_shape_cast(void* p, const pxg_cast_query* c, const pxg_query_filter* f, pxg_raycast_hit* out) {
    Handle* h = static_cast<Handle*>(p);
    std::memset(out, 0, sizeof(*out));
    int r = 0;
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        Shape<T> s = shape_from<T>(c->shape);
        RayHit<T> hit;
        bool started = false;
        if (!world_shape_cast(w, s, v3<T>(c->position), q4<T>(c->orientation), v3<T>(c->direction), T(c->max_distance),
                              filter_from(f), hit, started)) return;
        hit_to(hit, *out);
        out->started_inside = started ? 1u : 0u;
        r = 1;
    };
    if (h->dbl) go(h->d); else go(h->f);
    return r;
}
// compute_penetration(): every penetration in slot order; returns the count, copies min(cap, count)
int pxo_compute_penetration(void* p, const pxg_shape* shape, const float* pos, const float* rot,
                            const pxg_query_filter* f, pxg_penetration* out, int cap) {
    Handle* h = static_cast<Handle*>(p);
    int count = 0;
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        auto pens = world_penetration(w, shape_from<T>(*shape), v3<T>(pos), q4<T>(rot), filter_from(f));
        count = int(pens.size());
        for (int i = 0; i < count && i < cap; ++i) {
            pxg_penetration& o = out[i];
            std::memset(&o, 0, sizeof(o));
            put3(o.normal, pens[i].normal);
            o.depth = float(pens[i].depth);
            put3(o.point, pens[i].point);
            o.body = pens[i].body;
            o.shape_index = pens[i].shape;
        }
    };
    if (h->dbl) go(h->d); else go(h->f);
    return count;
}
// raycast_shape for one shape at a pose (float)
int pxo_ray_shape(const pxg_shape* s, const float* pos, const float* rot, const pxg_ray* ray, float* t, float* n3) {
    Shape<float> sh = shape_from<float>(*s);
    float tt = 0.0f;
    V3<float> nn;
    if (!raycast_shape(ray_from<float>(*ray), sh, v3<float>(pos), q4<float>(rot), tt, nn)) return 0;
    *t = tt;
    put3(n3, nn);
    return 1;
}
int pxo_ray_aabb(const pxg_ray* ray, const float* mn, const float* mx) {
    return ray_aabb(ray_from<float>(*ray), v3<float>(mn), v3<float>(mx)) ? 1 : 0;
}

// events(), clear_events() (world.h:744-748)
uint32_t pxo_get_events(void* p, uint32_t* out4, uint32_t cap) {
    Handle* h = static_cast<Handle*>(p);
    return h->dbl ? get_events_t(h->d, out4, cap) : get_events_t(h->f, out4, cap);
}
void pxo_clear_events(void* p) {
    Handle* h = static_cast<Handle*>(p);
    h->f.events.clear();
    h->d.events.clear();
}

uint32_t pxo_get_pairs(void* p, uint32_t* ab, uint32_t cap) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) -> uint32_t {
        uint32_t n = uint32_t(w.pairs.size());
        for (uint32_t i = 0; i < n && i < cap; ++i) { ab[2 * i] = w.pairs[i].first; ab[2 * i + 1] = w.pairs[i].second; }
        return n;
    };
    return h->dbl ? go(h->d) : go(h->f);
}
uint32_t pxo_get_manifolds(void* p, pxg_manifold* out, uint32_t cap) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) -> uint32_t {
        uint32_t n = uint32_t(w.manifolds.size());
        for (uint32_t i = 0; i < n && i < cap; ++i) manifold_to(w.manifolds[i], out[i]);
        return n;
    };
    return h->dbl ? go(h->d) : go(h->f);
}
void pxo_get_fat(void* p, float* out6, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        for (uint32_t i = 0; i < n && i < w.shapes.size(); ++i) {
            if (w.proxy[i] < 0) { for (int k = 0; k < 6; ++k) out6[6 * i + k] = 0; continue; }
            auto b = w.fat(i);
            put3(out6 + 6 * i, b.mn); put3(out6 + 6 * i + 3, b.mx);
        }
    };
    if (h->dbl) go(h->d); else go(h->f);
}
void pxo_set_fat(void* p, const float* in6, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        std::vector<AABB<T>> boxes(w.shapes.size());
        for (uint32_t i = 0; i < n && i < w.shapes.size(); ++i) { boxes[i].mn = v3<T>(in6 + 6 * i); boxes[i].mx = v3<T>(in6 + 6 * i + 3); }
        w.set_fat(boxes);
    };
    if (h->dbl) go(h->d); else go(h->f);
}
// Cache entries sorted by key.
uint32_t pxo_get_cache(void* p, uint64_t* keys, float* imp, uint32_t cap) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) -> uint32_t {
        std::map<uint64_t, Cached> sorted(w.cache.begin(), w.cache.end());
        uint32_t i = 0;
        for (auto& kv : sorted) {
            if (i < cap) { keys[i] = kv.first; imp[3 * i] = float(kv.second.n); imp[3 * i + 1] = float(kv.second.t1); imp[3 * i + 2] = float(kv.second.t2); }
            ++i;
        }
        return uint32_t(sorted.size());
    };
    return h->dbl ? go(h->d) : go(h->f);
}
void pxo_set_cache(void* p, const uint64_t* keys, const float* imp, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        w.cache.clear();
        for (uint32_t i = 0; i < n; ++i) { Cached c; c.n = imp[3 * i]; c.t1 = imp[3 * i + 1]; c.t2 = imp[3 * i + 2]; w.cache[keys[i]] = c; }
    };
    if (h->dbl) go(h->d); else go(h->f);
}
uint64_t pxo_state_hash(void* p) {
    Handle* h = static_cast<Handle*>(p);
    return h->dbl ? h->d.state_hash() : h->f.state_hash();
}
// out: [vel_passes, residual, predicted, steps]
void pxo_stats(void* p, double* out) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) { out[0] = w.vel_passes; out[1] = double(w.last_residual); out[2] = w.last_predicted; out[3] = double(w.steps); };
    if (h->dbl) go(h->d); else go(h->f);
}
uint32_t pxo_residual_trace(void* p, double* out, uint32_t cap) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) -> uint32_t {
        for (uint32_t i = 0; i < w.residual_trace.size() && i < cap; ++i) out[i] = double(w.residual_trace[i]);
        return uint32_t(w.residual_trace.size());
    };
    return h->dbl ? go(h->d) : go(h->f);
}
// The velocity residual of the given manifolds (impulses included) over the world's current
// bodies, in the given order, without solving anything; the manifolds replace the world's. The
// per-point terms go to the residual detail.
double pxo_eval_residual(void* p, const pxg_manifold* ms, uint32_t n, uint32_t use_rest, float hstep) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) -> double {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        w.manifolds.clear();
        for (uint32_t i = 0; i < n; ++i) w.manifolds.push_back(manifold_from<T>(ms[i]));
        std::vector<Manifold<T>*> order;
        for (auto& m : w.manifolds) order.push_back(&m);
        return double(w.residual(order, use_rest != 0, T(hstep), &w.residual_detail));
    };
    return h->dbl ? go(h->d) : go(h->f);
}

// Per-point residual terms after the last velocity pass of each following step (16 doubles per
// point, see World::residual_detail); on = 0 stops recording. Returns the point count.
void pxo_want_residual_detail(void* p, uint32_t on) {
    Handle* h = static_cast<Handle*>(p);
    h->f.want_residual_detail = on != 0; h->d.want_residual_detail = on != 0;
}
uint32_t pxo_residual_detail(void* p, double* out, uint32_t cap) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) -> uint32_t {
        for (uint32_t i = 0; i < w.residual_detail.size() && i < cap; ++i)
            for (int k = 0; k < 16; ++k) out[16 * i + k] = w.residual_detail[i][k];
        return uint32_t(w.residual_detail.size());
    };
    return h->dbl ? go(h->d) : go(h->f);
}

// ---- single-function entry points for known-answer tests (FP32) ----
static std::vector<Hull<float>> g_kat_hulls;  // the hulls the single-pair entry points' shapes index
void pxo_set_kat_hulls(const pxg_hull* in, uint32_t n) {
    g_kat_hulls.clear();
    for (uint32_t i = 0; i < n; ++i) g_kat_hulls.push_back(hull_from(in[i]));
}
static std::vector<Mesh> g_kat_meshes;  // the meshes the single-pair entry points' shapes index
int pxo_set_kat_meshes(const pxg_mesh_desc* in, uint32_t n) {
    g_kat_meshes.assign(n, Mesh{});
    for (uint32_t i = 0; i < n; ++i)
        if (!mesh_from(in[i], g_kat_meshes[i])) return 0;
    return 1;
}
// generate_mesh_contacts (mesh_contact.h:838-884) with the BVH (use_bvh 1) or every triangle (0):
// normal from the mesh toward the shape
int pxo_mesh_contacts(const pxg_shape* s, const float* pos, const float* rot, const pxg_shape* ms, const float* mpos,
                      const float* mrot, int use_bvh, pxg_manifold* out) {
    Shape<float> S = shape_from<float>(*s), M = shape_from<float>(*ms);
    resolve_hull(S, g_kat_hulls, s->capsule_axis);
    resolve_mesh(M, g_kat_meshes, ms->capsule_axis);
    Manifold<float> m;
    bool hit = mesh_contacts(S, v3<float>(pos), q4<float>(rot), M, v3<float>(mpos), q4<float>(mrot), m, use_bvh != 0);
    manifold_to(m, *out);
    return hit ? 1 : 0;
}
// the ascending candidate triangles of mesh_contact_candidates (mesh_contact.h:330-340)
uint32_t pxo_mesh_candidates(const pxg_shape* s, const float* pos, const float* rot, const pxg_shape* ms,
                             const float* mpos, const float* mrot, uint32_t* out, uint32_t cap) {
    Shape<float> S = shape_from<float>(*s), M = shape_from<float>(*ms);
    resolve_hull(S, g_kat_hulls, s->capsule_axis);
    resolve_mesh(M, g_kat_meshes, ms->capsule_axis);
    if (!M.mesh) return 0;
    std::vector<uint32_t> c;
    mesh_candidates(S, v3<float>(pos), q4<float>(rot), *static_cast<const Mesh*>(M.mesh),
                    mesh_world_transform(M, v3<float>(mpos), q4<float>(mrot)), c);
    for (size_t i = 0; i < c.size() && i < cap; ++i) out[i] = c[i];
    return uint32_t(c.size());
}
// glm::inverse of a 4x4 (column-major, 16 floats)
void pxo_mat4_inverse(const float* in, float* out) {
    M4<float> m;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) m.m[c][r] = in[4 * c + r];
    M4<float> i = m4_inverse(m);
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) out[4 * c + r] = i.m[c][r];
}
// closest_point_on_triangle and barycentric_ (mesh_contact.h:61-98, 247-258)
void pxo_closest_on_triangle(const float* p, const float* a, const float* b, const float* c, float* out, float* bary) {
    put3(out, closest_on_triangle(v3<float>(p), v3<float>(a), v3<float>(b), v3<float>(c)));
    put3(bary, barycentric(v3<float>(p), v3<float>(a), v3<float>(b), v3<float>(c)));
}
int pxo_generate_contacts(const pxg_shape* sa, const float* pa, const float* ra, const pxg_shape* sb, const float* pb,
                          const float* rb, int spec, pxg_manifold* out) {
    Manifold<float> m;
    Shape<float> A = shape_from<float>(*sa), B = shape_from<float>(*sb);
    resolve_hull(A, g_kat_hulls, sa->capsule_axis);
    resolve_hull(B, g_kat_hulls, sb->capsule_axis);
    resolve_mesh(A, g_kat_meshes, sa->capsule_axis);
    resolve_mesh(B, g_kat_meshes, sb->capsule_axis);
    bool hit = generate_contacts(0u, A, v3<float>(pa), q4<float>(ra), 1u, B, v3<float>(pb), q4<float>(rb), m, spec != 0);
    manifold_to(m, *out);
    return hit ? 1 : 0;
}
// the world bounds and CCD extent of a shape (hulls from the single-pair table)
void pxo_shape_bounds(const pxg_shape* s, const float* pos, const float* rot, float* out6, float* min_extent) {
    Shape<float> S = shape_from<float>(*s);
    resolve_hull(S, g_kat_hulls, s->capsule_axis);
    resolve_mesh(S, g_kat_meshes, s->capsule_axis);
    AABB<float> b = world_bounds(S, v3<float>(pos), q4<float>(rot));
    out6[0] = b.mn.x; out6[1] = b.mn.y; out6[2] = b.mn.z; out6[3] = b.mx.x; out6[4] = b.mx.y; out6[5] = b.mx.z;
    *min_extent = World<float>::min_extent(S);
}
int pxo_box_box(const float* ca, const float* hea, const float* qa, const float* cb, const float* heb, const float* qb, int spec,
                pxg_manifold* out) {
    OBB<float> A{v3<float>(ca), v3<float>(hea), q4<float>(qa)};
    OBB<float> B{v3<float>(cb), v3<float>(heb), q4<float>(qb)};
    Manifold<float> m;
    bool hit = box_box(A, B, m, spec != 0);
    manifold_to(m, *out);
    return hit ? 1 : 0;
}
// SAT only: returns hit, writes axis, penetration/separation, normal
int pxo_sat(const float* ca, const float* hea, const float* qa, const float* cb, const float* heb, const float* qb, int spec,
            int* axis, float* depth, float* normal, int* separated) {
    OBB<float> A{v3<float>(ca), v3<float>(hea), q4<float>(qa)};
    OBB<float> B{v3<float>(cb), v3<float>(heb), q4<float>(qb)};
    SatResult<float> r;
    bool hit = sat_obb_obb(A, B, r, spec != 0);
    *axis = r.axis; *depth = r.separated ? r.sep : r.pen; put3(normal, r.n); *separated = r.separated;
    return hit ? 1 : 0;
}
void pxo_seg_seg(const float* p1, const float* q1, const float* p2, const float* q2, float* c1, float* c2) {
    V3<float> a, b;
    closest_points_seg_seg(v3<float>(p1), v3<float>(q1), v3<float>(p2), v3<float>(q2), a, b);
    put3(c1, a); put3(c2, b);
}
void pxo_world_bounds(const pxg_shape* s, const float* pos, const float* rot, float* out6) {
    AABB<float> b = world_bounds(shape_from<float>(*s), v3<float>(pos), q4<float>(rot));
    put3(out6, b.mn); put3(out6 + 3, b.mx);
}
float pxo_combine(float a, float b, int ma, int mb) { return combine(a, b, ma, mb); }
void pxo_orthonormal_basis(const float* n, float* t1, float* t2) {
    V3<float> a, b;
    orthonormal_basis(v3<float>(n), a, b);
    put3(t1, a); put3(t2, b);
}
void pxo_quat_rotate(const float* q, const float* v, float* out) { put3(out, qrot(q4<float>(q), v3<float>(v))); }
void pxo_mat3_cast(const float* q, float* out9) {
    M3<float> m = mat3_cast(q4<float>(q));
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) out9[c * 3 + r] = m.c[c][r];
}
uint64_t pxo_pack_key(uint32_t a, uint32_t b, uint32_t fid) { return pack_key(a, b, fid); }
// joint helpers (FP32, libm atan2 unless mode == 1)
void pxo_m3_inverse(const float* in9, float* out9) {
    M3<float> m;
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) m.c[c][r] = in9[c * 3 + r];
    M3<float> i = m3_inverse(m);
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) out9[c * 3 + r] = i.c[c][r];
}
float pxo_m3_det(const float* in9) {
    M3<float> m;
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) m.c[c][r] = in9[c * 3 + r];
    return m3_det(m);
}
int pxo_point_block_mass(const float* ra, const float* rb, float ima, float imb, const float* Ia9, const float* Ib9, float* out9) {
    M3<float> A, B, K;
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) { A.c[c][r] = Ia9[c * 3 + r]; B.c[c][r] = Ib9[c * 3 + r]; }
    int ok = point_block_mass(v3<float>(ra), v3<float>(rb), ima, imb, A, B, K) ? 1 : 0;
    for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) out9[c * 3 + r] = K.c[c][r];
    return ok;
}
float pxo_hinge_angle(const float* qa, const float* qb, const float* axis, const float* rest, uint32_t mode) {
    return hinge_angle(q4<float>(qa), q4<float>(qb), v3<float>(axis), q4<float>(rest), mode ? Atan2Mode::Rounded : Atan2Mode::Libm);
}
float pxo_cone_swing(const float* ta, const float* tb, float* axis, uint32_t mode) {
    V3<float> s;
    float a = cone_swing(v3<float>(ta), v3<float>(tb), s, mode ? Atan2Mode::Rounded : Atan2Mode::Libm);
    put3(axis, s);
    return a;
}
float pxo_atan2(float y, float x, uint32_t mode) { return atan2_of(y, x, mode ? Atan2Mode::Rounded : Atan2Mode::Libm); }


// ---- cloth and hair ----
}  // extern "C"
namespace {
template <class T> CParticle<T> particle_from(const pxg_particle& d) {
    CParticle<T> p;
    p.pos = v3<T>(d.position); p.prev = v3<T>(d.prev_position); p.vel = v3<T>(d.velocity); p.inv_mass = T(d.inv_mass);
    return p;
}
template <class T> void particle_to(const CParticle<T>& p, pxg_particle& d) {
    std::memset(&d, 0, sizeof(d));
    put3(d.position, p.pos); put3(d.prev_position, p.prev); put3(d.velocity, p.vel); d.inv_mass = float(p.inv_mass);
}
template <class T> std::vector<CConstraint<T>> constraints_from(const pxg_distance_constraint* c, uint32_t n) {
    std::vector<CConstraint<T>> out(n);
    for (uint32_t i = 0; i < n; ++i) { out[i].a = c[i].a; out[i].b = c[i].b; out[i].rest = T(c[i].rest_length); out[i].lambda = T(c[i].lambda); }
    return out;
}
std::vector<CBatch> batches_from(const uint32_t* b, uint32_t n) {
    std::vector<CBatch> out(n);
    for (uint32_t i = 0; i < n; ++i) out[i] = CBatch{b[2 * i], b[2 * i + 1]};
    return out;
}
template <class T> ClothParams<T> cloth_params_from(const pxg_cloth_params& p) {
    ClothParams<T> o;
    o.stretch_compliance = T(p.stretch_compliance); o.bend_compliance = T(p.bend_compliance); o.damping = T(p.damping);
    o.thickness = T(p.thickness); o.friction = T(p.friction); o.gravity_scale = T(p.gravity_scale);
    o.max_velocity = T(p.max_velocity); o.external_acceleration = v3<T>(p.external_acceleration); o.wind = v3<T>(p.wind);
    o.wind_turbulence = T(p.wind_turbulence); o.air_drag = T(p.air_drag); o.air_lift = T(p.air_lift);
    o.self_collision = p.self_collision != 0; o.self_distance = T(p.self_distance); o.tether_scale = T(p.tether_scale);
    o.substeps = p.substeps; o.iterations = p.iterations; o.sleep_threshold = T(p.sleep_threshold);
    o.sleep_time = T(p.sleep_time); o.layer = p.layer;
    return o;
}
template <class T> void cloth_params_to(const ClothParams<T>& o, pxg_cloth_params& p) {
    p.stretch_compliance = float(o.stretch_compliance); p.bend_compliance = float(o.bend_compliance); p.damping = float(o.damping);
    p.thickness = float(o.thickness); p.friction = float(o.friction); p.gravity_scale = float(o.gravity_scale);
    p.max_velocity = float(o.max_velocity); put3(p.external_acceleration, o.external_acceleration); put3(p.wind, o.wind);
    p.wind_turbulence = float(o.wind_turbulence); p.air_drag = float(o.air_drag); p.air_lift = float(o.air_lift);
    p.self_collision = o.self_collision; p.self_distance = float(o.self_distance); p.tether_scale = float(o.tether_scale);
    p.substeps = o.substeps; p.iterations = o.iterations; p.sleep_threshold = float(o.sleep_threshold);
    p.sleep_time = float(o.sleep_time); p.layer = o.layer;
}
template <class T> Cloth<T> cloth_from(const pxg_cloth& d) {
    Cloth<T> c;
    for (uint32_t i = 0; i < d.n_particles; ++i) c.particles.push_back(particle_from<T>(d.particles[i]));
    c.columns = d.columns; c.rows = d.rows;
    c.stretch = constraints_from<T>(d.stretch, d.n_stretch);
    c.bend = constraints_from<T>(d.bend, d.n_bend);
    c.stretch_batches = batches_from(d.stretch_batches, d.n_stretch_batches);
    c.bend_batches = batches_from(d.bend_batches, d.n_bend_batches);
    for (uint32_t i = 0; i < d.n_tethers; ++i) c.tethers.push_back(CTether<T>{d.tethers[i].particle, d.tethers[i].anchor, T(d.tethers[i].max_length)});
    for (uint32_t i = 0; i < d.n_anchors; ++i) {
        const pxg_cloth_anchor& a = d.anchors[i];
        CAnchor<T> o;
        o.particle = a.particle; o.has_body = a.has_body != 0; o.body = a.body; o.generation = a.generation;
        o.local = v3<T>(a.local_position); o.world = v3<T>(a.world_position); o.prev_world = v3<T>(a.prev_world_position);
        c.anchors.push_back(o);
    }
    c.triangles.assign(d.triangles, d.triangles + d.n_triangle_indices);
    c.params = cloth_params_from<T>(d.params);
    c.enabled = d.enabled != 0; c.awake = d.awake != 0; c.sleep_timer = T(d.sleep_timer); c.time = T(d.time);
    c.bounds.mn = v3<T>(d.bounds_min); c.bounds.mx = v3<T>(d.bounds_max);
    return c;
}
// writes into the arrays of `d` (sizes must match what was uploaded)
template <class T> void cloth_to(const Cloth<T>& c, pxg_cloth& d) {
    for (size_t i = 0; d.particles && i < c.particles.size() && i < d.n_particles; ++i) particle_to(c.particles[i], d.particles[i]);
    for (size_t i = 0; d.stretch && i < c.stretch.size() && i < d.n_stretch; ++i) d.stretch[i].lambda = float(c.stretch[i].lambda);
    for (size_t i = 0; d.bend && i < c.bend.size() && i < d.n_bend; ++i) d.bend[i].lambda = float(c.bend[i].lambda);
    for (size_t i = 0; d.anchors && i < c.anchors.size() && i < d.n_anchors; ++i) {
        put3(d.anchors[i].world_position, c.anchors[i].world);
        put3(d.anchors[i].prev_world_position, c.anchors[i].prev_world);
    }
    d.awake = c.awake; d.enabled = c.enabled; d.sleep_timer = float(c.sleep_timer); d.time = float(c.time);
    put3(d.bounds_min, c.bounds.mn); put3(d.bounds_max, c.bounds.mx);
}
template <class T> StrandParams<T> strand_params_from(const pxg_strand_params& p) {
    StrandParams<T> o;
    o.stretch_compliance = T(p.stretch_compliance); o.global_shape = T(p.global_shape);
    o.global_shape_falloff = T(p.global_shape_falloff); o.local_shape = T(p.local_shape); o.damping = T(p.damping);
    o.thickness = T(p.thickness); o.friction = T(p.friction); o.gravity_scale = T(p.gravity_scale);
    o.max_velocity = T(p.max_velocity); o.external_acceleration = v3<T>(p.external_acceleration); o.wind = v3<T>(p.wind);
    o.wind_turbulence = T(p.wind_turbulence); o.air_drag = T(p.air_drag); o.substeps = p.substeps;
    o.iterations = p.iterations; o.sleep_threshold = T(p.sleep_threshold); o.sleep_time = T(p.sleep_time); o.layer = p.layer;
    return o;
}
template <class T> StrandSet<T> strands_from(const pxg_strand_set& d) {
    StrandSet<T> s;
    for (uint32_t i = 0; i < d.n_particles; ++i) {
        s.particles.push_back(particle_from<T>(d.particles[i]));
        s.rest_local.push_back(v3<T>(d.rest_local + 3 * i));
        s.rest_offset.push_back(v3<T>(d.rest_offset + 3 * i));
    }
    s.stretch = constraints_from<T>(d.stretch, d.n_stretch);
    s.stretch_batches = batches_from(d.stretch_batches, d.n_stretch_batches);
    for (uint32_t i = 0; i < d.n_strands; ++i) {
        const pxg_strand& x = d.strands[i];
        Strand<T> st;
        st.first = x.first; st.count = x.count; st.root = v3<T>(x.root_position); st.root_q = q4<T>(x.root_orientation);
        st.prev_root = v3<T>(x.prev_root_position); st.prev_root_q = q4<T>(x.prev_root_orientation);
        s.strands.push_back(st);
    }
    for (uint32_t i = 0; i < d.n_collide_bodies; ++i) s.collide_bodies.push_back({d.collide_bodies[2 * i], d.collide_bodies[2 * i + 1]});
    s.params = strand_params_from<T>(d.params);
    s.enabled = d.enabled != 0; s.awake = d.awake != 0; s.sleep_timer = T(d.sleep_timer); s.time = T(d.time);
    s.bounds.mn = v3<T>(d.bounds_min); s.bounds.mx = v3<T>(d.bounds_max);
    return s;
}
template <class T> void strands_to(const StrandSet<T>& s, pxg_strand_set& d) {
    for (size_t i = 0; d.particles && i < s.particles.size() && i < d.n_particles; ++i) particle_to(s.particles[i], d.particles[i]);
    for (size_t i = 0; d.stretch && i < s.stretch.size() && i < d.n_stretch; ++i) d.stretch[i].lambda = float(s.stretch[i].lambda);
    for (size_t i = 0; d.strands && i < s.strands.size() && i < d.n_strands; ++i) {
        put3(d.strands[i].root_position, s.strands[i].root); put4(d.strands[i].root_orientation, s.strands[i].root_q);
        put3(d.strands[i].prev_root_position, s.strands[i].prev_root); put4(d.strands[i].prev_root_orientation, s.strands[i].prev_root_q);
    }
    d.awake = s.awake; d.enabled = s.enabled; d.sleep_timer = float(s.sleep_timer); d.time = float(s.time);
    put3(d.bounds_min, s.bounds.mn); put3(d.bounds_max, s.bounds.mx);
}
}  // namespace
extern "C" {

void pxo_set_cloths(void* p, const pxg_cloth* in, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        w.cloths.clear(); w.cloth_alive.clear();
        for (uint32_t i = 0; i < n; ++i) {
            w.cloths.push_back(cloth_from<T>(in[i]));
            w.cloths.back().bounds = particle_bounds(w.cloths.back().particles);  // add_cloth, world.h:555
            w.cloth_alive.push_back(in[i].alive && in[i].n_particles ? 1 : 0);
        }
    };
    go(h->f); go(h->d);
}
void pxo_get_cloth(void* p, uint32_t index, pxg_cloth* out) {
    Handle* h = static_cast<Handle*>(p);
    if (h->dbl) { if (index < h->d.cloths.size()) cloth_to(h->d.cloths[index], *out); }
    else if (index < h->f.cloths.size()) cloth_to(h->f.cloths[index], *out);
}
void pxo_set_strand_sets(void* p, const pxg_strand_set* in, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        w.strand_sets.clear(); w.strand_alive.clear();
        for (uint32_t i = 0; i < n; ++i) {
            w.strand_sets.push_back(strands_from<T>(in[i]));
            w.strand_sets.back().bounds = particle_bounds(w.strand_sets.back().particles);  // add_strands, world.h:638
            w.strand_alive.push_back(in[i].alive && in[i].n_particles && in[i].n_strands ? 1 : 0);
        }
    };
    go(h->f); go(h->d);
}
void pxo_get_strand_set(void* p, uint32_t index, pxg_strand_set* out) {
    Handle* h = static_cast<Handle*>(p);
    if (h->dbl) { if (index < h->d.strand_sets.size()) strands_to(h->d.strand_sets[index], *out); }
    else if (index < h->f.strand_sets.size()) strands_to(h->f.strand_sets[index], *out);
}
void pxo_set_strand_roots(void* p, uint32_t index, const float* pos, const float* rot, uint32_t n) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& w) {
        using T = std::remove_reference_t<decltype(w.bodies[0].mass)>;
        if (index >= w.strand_sets.size()) return;
        std::vector<V3<T>> ps(n);
        std::vector<Q<T>> qs(rot ? n : 0);
        for (uint32_t i = 0; i < n; ++i) ps[i] = v3<T>(pos + 3 * i);
        for (uint32_t i = 0; rot && i < n; ++i) qs[i] = q4<T>(rot + 4 * i);
        set_strand_roots(w.strand_sets[index], ps, qs);
    };
    go(h->f); go(h->d);
}
// step() (world.h:1229-1257): returns the fixed substeps run
uint32_t pxo_step_frame(void* p, float dt) {
    Handle* h = static_cast<Handle*>(p);
    return h->dbl ? h->d.step(double(dt)) : h->f.step(dt);
}
float pxo_interpolation_alpha(void* p) {
    Handle* h = static_cast<Handle*>(p);
    return h->dbl ? float(h->d.interpolation_alpha) : h->f.interpolation_alpha;
}
// 1: colliders in the reference's tree order; 0: in shape slot order within each tree
void pxo_set_collider_order(void* p, uint32_t tree) {
    Handle* h = static_cast<Handle*>(p);
    h->f.collider_tree_order = tree != 0;
    h->d.collider_tree_order = tree != 0;
}

// ---- builders (test fixtures): grid cloth with pins and tethers, strands ----
struct ClothBuild { Cloth<float> c; };
// fd: width, height, center xyz, orientation xyzw, total_mass; ui: columns, rows, shear
void* pxo_cloth_grid(const float* fd, const uint32_t* ui, const pxg_cloth_params* params) {
    GridClothDesc<float> d;
    d.width = fd[0]; d.height = fd[1]; d.center = v3<float>(fd + 2); d.orientation = q4<float>(fd + 5); d.total_mass = fd[9];
    d.columns = ui[0]; d.rows = ui[1]; d.shear = ui[2] != 0;
    d.params = cloth_params_from<float>(*params);
    return new ClothBuild{make_grid_cloth(d)};
}
// pin_static (body == nullptr) or pin_to_body; body_pose: position xyz, orientation xyzw
uint32_t pxo_cloth_pin(void* b, const float* point, float radius, const uint32_t* body, const float* body_pose) {
    ClothBuild* cb = static_cast<ClothBuild*>(b);
    if (!body) return pin_particles(cb->c, v3<float>(point), radius, false, 0u, 0u, V3<float>(), Q<float>());
    return pin_particles(cb->c, v3<float>(point), radius, true, body[0], body[1], v3<float>(body_pose), q4<float>(body_pose + 3));
}
void pxo_cloth_tethers(void* b) { build_tethers(static_cast<ClothBuild*>(b)->c); }
// sizes: particles, stretch, bend, stretch batches, bend batches, tethers, anchors, triangle indices
void pxo_cloth_sizes(void* b, uint32_t* s) {
    const Cloth<float>& c = static_cast<ClothBuild*>(b)->c;
    s[0] = uint32_t(c.particles.size()); s[1] = uint32_t(c.stretch.size()); s[2] = uint32_t(c.bend.size());
    s[3] = uint32_t(c.stretch_batches.size()); s[4] = uint32_t(c.bend_batches.size()); s[5] = uint32_t(c.tethers.size());
    s[6] = uint32_t(c.anchors.size()); s[7] = uint32_t(c.triangles.size());
}
void pxo_cloth_export(void* b, pxg_cloth* d, uint32_t* batches_out) {
    const Cloth<float>& c = static_cast<ClothBuild*>(b)->c;
    for (size_t i = 0; i < c.particles.size(); ++i) particle_to(c.particles[i], d->particles[i]);
    for (size_t i = 0; i < c.stretch.size(); ++i) d->stretch[i] = pxg_distance_constraint{c.stretch[i].a, c.stretch[i].b, c.stretch[i].rest, 0.0f};
    for (size_t i = 0; i < c.bend.size(); ++i) d->bend[i] = pxg_distance_constraint{c.bend[i].a, c.bend[i].b, c.bend[i].rest, 0.0f};
    uint32_t k = 0;
    for (const CBatch& x : c.stretch_batches) { batches_out[k++] = x.begin; batches_out[k++] = x.end; }
    for (const CBatch& x : c.bend_batches) { batches_out[k++] = x.begin; batches_out[k++] = x.end; }
    pxg_cloth_tether* t = const_cast<pxg_cloth_tether*>(d->tethers);
    for (size_t i = 0; i < c.tethers.size(); ++i) t[i] = pxg_cloth_tether{c.tethers[i].particle, c.tethers[i].anchor, c.tethers[i].max_len, 0u};
    for (size_t i = 0; i < c.anchors.size(); ++i) {
        pxg_cloth_anchor& a = d->anchors[i];
        std::memset(&a, 0, sizeof(a));
        a.particle = c.anchors[i].particle; a.has_body = c.anchors[i].has_body; a.body = c.anchors[i].body;
        a.generation = c.anchors[i].generation; put3(a.local_position, c.anchors[i].local);
        put3(a.world_position, c.anchors[i].world); put3(a.prev_world_position, c.anchors[i].prev_world);
    }
    uint32_t* tri = const_cast<uint32_t*>(d->triangles);
    for (size_t i = 0; i < c.triangles.size(); ++i) tri[i] = c.triangles[i];
    d->columns = c.columns; d->rows = c.rows;
    cloth_params_to(c.params, d->params);
    put3(d->bounds_min, c.bounds.mn); put3(d->bounds_max, c.bounds.mx);
}
void pxo_cloth_free(void* b) { delete static_cast<ClothBuild*>(b); }

// make_strands: counts[i] points per strand (xyz in pts), root orientations (xyzw) and masses per strand
struct StrandBuild { StrandSet<float> s; };
void* pxo_strands_make(const float* pts, const uint32_t* counts, uint32_t n_strands, const float* root_q,
                       const float* masses, const pxg_strand_params* params) {
    std::vector<StrandDesc<float>> ds(n_strands);
    size_t k = 0;
    for (uint32_t i = 0; i < n_strands; ++i) {
        for (uint32_t j = 0; j < counts[i]; ++j, ++k) ds[i].points.push_back(v3<float>(pts + 3 * k));
        ds[i].root_q = q4<float>(root_q + 4 * i);
        ds[i].particle_mass = masses[i];
    }
    return new StrandBuild{make_strands(ds, strand_params_from<float>(*params))};
}
// sizes: particles, stretch, batches, strands
void pxo_strands_sizes(void* b, uint32_t* s) {
    const StrandSet<float>& x = static_cast<StrandBuild*>(b)->s;
    s[0] = uint32_t(x.particles.size()); s[1] = uint32_t(x.stretch.size()); s[2] = uint32_t(x.stretch_batches.size());
    s[3] = uint32_t(x.strands.size());
}
void pxo_strands_export(void* b, pxg_strand_set* d) {
    const StrandSet<float>& x = static_cast<StrandBuild*>(b)->s;
    float* rl = const_cast<float*>(d->rest_local);
    float* ro = const_cast<float*>(d->rest_offset);
    for (size_t i = 0; i < x.particles.size(); ++i) {
        particle_to(x.particles[i], d->particles[i]);
        put3(rl + 3 * i, x.rest_local[i]);
        put3(ro + 3 * i, x.rest_offset[i]);
    }
    for (size_t i = 0; i < x.stretch.size(); ++i) d->stretch[i] = pxg_distance_constraint{x.stretch[i].a, x.stretch[i].b, x.stretch[i].rest, 0.0f};
    uint32_t* bt = const_cast<uint32_t*>(d->stretch_batches);
    for (size_t i = 0; i < x.stretch_batches.size(); ++i) { bt[2 * i] = x.stretch_batches[i].begin; bt[2 * i + 1] = x.stretch_batches[i].end; }
    for (size_t i = 0; i < x.strands.size(); ++i) {
        pxg_strand& st = d->strands[i];
        st.first = x.strands[i].first; st.count = x.strands[i].count;
        put3(st.root_position, x.strands[i].root); put4(st.root_orientation, x.strands[i].root_q);
        put3(st.prev_root_position, x.strands[i].prev_root); put4(st.prev_root_orientation, x.strands[i].prev_root_q);
    }
    put3(d->bounds_min, x.bounds.mn); put3(d->bounds_max, x.bounds.mx);
}
void pxo_strands_free(void* b) { delete static_cast<StrandBuild*>(b); }

}  // extern "C"
