// Cloth and hair records and builders for the C++ facade (include/pxg_world.hpp), over the ABI
// records of pxg.h: cloth::Cloth / hair::StrandSet as host vectors (cloth/cloth.h:249-315,
// hair/strand.h:187-240), and the reference's builders (cloth/cloth_builder.h: make_grid_cloth,
// pin_to_body, pin_static, build_tethers; hair/strand_builder.h: make_strands,
// set_strand_roots) with the same steps, so a sheet or groom built here is the one the reference
// builds. Host code only; it follows glm's evaluation order and needs a compiler that does not
// contract a * b + c into a fused multiply-add (-ffp-contract=off) to reproduce the reference's
// floats bit for bit.
#pragma once

#include "pxg.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <queue>
#include <utility>
#include <vector>

namespace pxg {

// Generational handles of the cloth and hair tables (world.h:552-700).
struct ClothId {
    uint32_t index = 0xFFFFFFFFu;
    uint32_t generation = 0;
    bool is_valid() const { return index != 0xFFFFFFFFu; }
    bool operator==(const ClothId& o) const { return index == o.index && generation == o.generation; }
};
struct StrandSetId {
    uint32_t index = 0xFFFFFFFFu;
    uint32_t generation = 0;
    bool is_valid() const { return index != 0xFFFFFFFFu; }
    bool operator==(const StrandSetId& o) const { return index == o.index && generation == o.generation; }
};

// cloth::ClothParams defaults (cloth/cloth.h:120-226)
inline pxg_cloth_params default_cloth_params() {
    pxg_cloth_params p;
    std::memset(&p, 0, sizeof(p));
    p.bend_compliance = 5e-5f;
    p.damping = 0.05f;
    p.thickness = 0.03f;
    p.friction = 0.5f;
    p.gravity_scale = 1.0f;
    p.max_velocity = 100.0f;
    p.air_drag = 0.02f;
    p.air_lift = 0.02f;
    p.tether_scale = 1.02f;
    p.sleep_threshold = 0.05f;
    p.sleep_time = 0.5f;
    return p;
}
// hair::StrandParams defaults (hair/strand.h:81-180)
inline pxg_strand_params default_strand_params() {
    pxg_strand_params p;
    std::memset(&p, 0, sizeof(p));
    p.global_shape = 0.22f;
    p.global_shape_falloff = 1.6f;
    p.local_shape = 0.35f;
    p.damping = 0.08f;
    p.thickness = 0.01f;
    p.friction = 0.4f;
    p.gravity_scale = 1.0f;
    p.max_velocity = 50.0f;
    p.wind_turbulence = 0.1f;
    p.air_drag = 0.6f;
    p.sleep_threshold = 0.02f;
    p.sleep_time = 0.6f;
    return p;
}

// cloth::Cloth: particles, stretch and bend constraints with their particle-disjoint batches
// ([begin, end) pairs), tethers, anchors, the triangles aerodynamics uses, the grid size
// (columns 0: not a grid) and the sleep state.
struct Cloth {
    std::vector<pxg_particle> particles;
    std::vector<pxg_distance_constraint> stretch, bend;
    std::vector<uint32_t> stretch_batches, bend_batches;
    std::vector<pxg_cloth_tether> tethers;
    std::vector<pxg_cloth_anchor> anchors;
    std::vector<uint32_t> triangles;
    uint32_t columns = 0, rows = 0;
    pxg_cloth_params params = default_cloth_params();
    bool enabled = true, awake = true;
    float sleep_timer = 0.0f, time = 0.0f;
    float bounds_min[3] = {0.0f, 0.0f, 0.0f}, bounds_max[3] = {0.0f, 0.0f, 0.0f};
};

// hair::StrandSet: particles, stretch constraints and batches, rest poses (3 floats per
// particle), strands, the bodies it collides with (all when empty) and the sleep state.
struct StrandSet {
    std::vector<pxg_particle> particles;
    std::vector<pxg_distance_constraint> stretch;
    std::vector<uint32_t> stretch_batches;
    std::vector<float> rest_local, rest_offset;
    std::vector<pxg_strand> strands;
    std::vector<std::pair<uint32_t, uint32_t>> collide_bodies;  // (body index, generation)
    pxg_strand_params params = default_strand_params();
    bool enabled = true, awake = true;
    float sleep_timer = 0.0f, time = 0.0f;
    float bounds_min[3] = {0.0f, 0.0f, 0.0f}, bounds_max[3] = {0.0f, 0.0f, 0.0f};
};

namespace soft_math {  // glm 0.9.9's evaluation order
struct V {
    float x, y, z;
};
inline V v(const float* p) { return V{p[0], p[1], p[2]}; }
inline void put(float* d, V a) { d[0] = a.x; d[1] = a.y; d[2] = a.z; }
inline V add(V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V sub(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V mul(V a, float s) { return V{a.x * s, a.y * s, a.z * s}; }
inline float dot(V a, V b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline float length(V a) { return std::sqrt(dot(a, a)); }
inline V cross(V a, V b) { return V{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
// q (x, y, z, w) * v: v + ((qv x v) * w + qv x (qv x v)) * 2
inline V rotate(const float* q, V a) {
    const V qv{q[0], q[1], q[2]};
    const V uv = cross(qv, a), uuv = cross(qv, uv);
    return add(a, mul(add(mul(uv, q[3]), uuv), 2.0f));
}
// glm::inverse(quat): conjugate / dot(q, q)
inline void inverse(const float* q, float* o) {
    const float d = (q[0] * q[0] + q[1] * q[1]) + (q[2] * q[2] + q[3] * q[3]);
    o[0] = -q[0] / d;
    o[1] = -q[1] / d;
    o[2] = -q[2] / d;
    o[3] = q[3] / d;
}
}  // namespace soft_math

// cloth::GridClothDesc (cloth/cloth_builder.h:52-64); orientation x, y, z, w
struct GridClothDesc {
    uint32_t columns = 21, rows = 21;
    float width = 2.0f, height = 2.0f;
    float center[3] = {0.0f, 0.0f, 0.0f};
    float orientation[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    float total_mass = 1.0f;
    bool shear = true;
    pxg_cloth_params params = default_cloth_params();
};

// cloth::compute_bounds (cloth/cloth.h:320-327)
inline void compute_bounds(const std::vector<pxg_particle>& ps, float* mn, float* mx) {
    for (int k = 0; k < 3; ++k) {
        mn[k] = std::numeric_limits<float>::max();
        mx[k] = -std::numeric_limits<float>::max();
    }
    for (const pxg_particle& p : ps)
        for (int k = 0; k < 3; ++k) {
            if (p.position[k] < mn[k]) mn[k] = p.position[k];
            if (mx[k] < p.position[k]) mx[k] = p.position[k];
        }
}

// make_grid_cloth (cloth/cloth_builder.h:81-219): row-major particles, structural edges in
// parity batches along x then y, optional shear diagonals (both directions, parity by column),
// bend edges two apart in four phases per axis, two triangles per cell.
inline Cloth make_grid_cloth(const GridClothDesc& d) {
    using namespace soft_math;
    Cloth c;
    const uint32_t cols = std::max<uint32_t>(2, d.columns), rows = std::max<uint32_t>(2, d.rows);
    c.columns = cols;
    c.rows = rows;
    c.params = d.params;
    const uint32_t count = cols * rows;
    const float per = d.total_mass > 1e-6f ? d.total_mass / float(count) : 0.0f;
    const float inv = per > 1e-6f ? 1.0f / per : 0.0f;
    const float dx = d.width / float(cols - 1), dy = d.height / float(rows - 1);
    c.particles.assign(count, pxg_particle{});
    for (uint32_t y = 0; y < rows; ++y)
        for (uint32_t x = 0; x < cols; ++x) {
            const V local{float(x) * dx - d.width * 0.5f, float(y) * dy - d.height * 0.5f, 0.0f};
            pxg_particle& p = c.particles[y * cols + x];
            std::memset(&p, 0, sizeof(p));
            const V w = add(v(d.center), rotate(d.orientation, local));
            put(p.position, w);
            put(p.prev_position, w);
            p.inv_mass = inv;
        }
    auto at = [cols](uint32_t x, uint32_t y) { return y * cols + x; };
    auto edge = [&c](std::vector<pxg_distance_constraint>& out, uint32_t a, uint32_t b) {
        pxg_distance_constraint k{a, b, length(sub(v(c.particles[a].position), v(c.particles[b].position))), 0.0f};
        out.push_back(k);
    };
    auto batch = [](const std::vector<pxg_distance_constraint>& list, std::vector<uint32_t>& batches, uint32_t begin) {
        const uint32_t end = static_cast<uint32_t>(list.size());
        if (end > begin) {
            batches.push_back(begin);
            batches.push_back(end);
        }
    };
    for (uint32_t par = 0; par < 2; ++par) {  // structural along x
        const uint32_t b = static_cast<uint32_t>(c.stretch.size());
        for (uint32_t y = 0; y < rows; ++y)
            for (uint32_t x = par; x + 1 < cols; x += 2) edge(c.stretch, at(x, y), at(x + 1, y));
        batch(c.stretch, c.stretch_batches, b);
    }
    for (uint32_t par = 0; par < 2; ++par) {  // structural along y
        const uint32_t b = static_cast<uint32_t>(c.stretch.size());
        for (uint32_t y = par; y + 1 < rows; y += 2)
            for (uint32_t x = 0; x < cols; ++x) edge(c.stretch, at(x, y), at(x, y + 1));
        batch(c.stretch, c.stretch_batches, b);
    }
    if (d.shear) {
        for (uint32_t par = 0; par < 2; ++par) {
            const uint32_t b = static_cast<uint32_t>(c.stretch.size());
            for (uint32_t y = 0; y + 1 < rows; ++y)
                for (uint32_t x = par; x + 1 < cols; x += 2) edge(c.stretch, at(x, y), at(x + 1, y + 1));
            batch(c.stretch, c.stretch_batches, b);
        }
        for (uint32_t par = 0; par < 2; ++par) {
            const uint32_t b = static_cast<uint32_t>(c.stretch.size());
            for (uint32_t y = 0; y + 1 < rows; ++y)
                for (uint32_t x = par; x + 1 < cols; x += 2) edge(c.stretch, at(x + 1, y), at(x, y + 1));
            batch(c.stretch, c.stretch_batches, b);
        }
    }
    for (uint32_t ph = 0; ph < 4; ++ph) {  // bend along x
        const uint32_t b = static_cast<uint32_t>(c.bend.size());
        for (uint32_t y = 0; y < rows; ++y)
            for (uint32_t x = ph; x + 2 < cols; x += 4) edge(c.bend, at(x, y), at(x + 2, y));
        batch(c.bend, c.bend_batches, b);
    }
    for (uint32_t ph = 0; ph < 4; ++ph) {  // bend along y
        const uint32_t b = static_cast<uint32_t>(c.bend.size());
        for (uint32_t y = ph; y + 2 < rows; y += 4)
            for (uint32_t x = 0; x < cols; ++x) edge(c.bend, at(x, y), at(x, y + 2));
        batch(c.bend, c.bend_batches, b);
    }
    for (uint32_t y = 0; y + 1 < rows; ++y)
        for (uint32_t x = 0; x + 1 < cols; ++x) {
            const uint32_t i00 = at(x, y), i10 = at(x + 1, y), i01 = at(x, y + 1), i11 = at(x + 1, y + 1);
            c.triangles.insert(c.triangles.end(), {i00, i10, i11, i00, i11, i01});
        }
    if (c.params.self_distance <= 0.0f) c.params.self_distance = 0.6f * std::min(dx, dy);
    compute_bounds(c.particles, c.bounds_min, c.bounds_max);
    return c;
}

namespace detail {
// pin_to_body / pin_static (cloth/cloth_builder.h:235-292): every free particle within radius
// becomes a pinned anchor, following the body from its current pose or fixed in place
inline uint32_t pin(Cloth& c, const float* point, float radius, bool has_body, uint32_t body, uint32_t generation,
                    const float* body_pos, const float* body_rot) {
    using namespace soft_math;
    const float r2 = radius * radius;
    float inv[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    if (has_body) inverse(body_rot, inv);
    uint32_t pinned = 0;
    for (uint32_t i = 0; i < c.particles.size(); ++i) {
        pxg_particle& p = c.particles[i];
        if (p.inv_mass == 0.0f) continue;
        const V d = sub(v(p.position), v(point));
        if (dot(d, d) > r2) continue;
        p.inv_mass = 0.0f;
        p.velocity[0] = p.velocity[1] = p.velocity[2] = 0.0f;
        pxg_cloth_anchor a;
        std::memset(&a, 0, sizeof(a));
        a.particle = i;
        a.has_body = has_body ? 1u : 0u;
        a.body = body;
        a.generation = generation;
        if (has_body) put(a.local_position, rotate(inv, sub(v(p.position), v(body_pos))));
        std::memcpy(a.world_position, p.position, sizeof(a.world_position));
        std::memcpy(a.prev_world_position, p.position, sizeof(a.prev_world_position));
        c.anchors.push_back(a);
        ++pinned;
    }
    return pinned;
}
}  // namespace detail

inline uint32_t pin_static(Cloth& c, const float world_point[3], float radius) {
    return detail::pin(c, world_point, radius, false, 0u, 0u, nullptr, nullptr);
}
// body: index and generation of the BodyId to follow; its current position and orientation
inline uint32_t pin_to_body(Cloth& c, uint32_t body, uint32_t generation, const pxg_body& pose, const float world_point[3],
                            float radius) {
    return detail::pin(c, world_point, radius, true, body, generation, pose.position, pose.orientation);
}

// build_tethers (cloth/cloth_builder.h:308-359): Dijkstra over the stretch graph from every
// anchor; each free particle reached is held within tether_scale x its path length of the
// nearest anchor.
inline void build_tethers(Cloth& c) {
    using namespace soft_math;
    c.tethers.clear();
    if (c.anchors.empty() || c.particles.empty()) return;
    const size_t n = c.particles.size();
    std::vector<std::vector<uint32_t>> adj(n);
    for (const pxg_distance_constraint& k : c.stretch) {
        adj[k.a].push_back(k.b);
        adj[k.b].push_back(k.a);
    }
    const float inf = std::numeric_limits<float>::max();
    std::vector<float> dist(n, inf);
    std::vector<uint32_t> from(n, 0u);
    using Entry = std::pair<float, uint32_t>;
    std::priority_queue<Entry, std::vector<Entry>, std::greater<Entry>> open;
    for (const pxg_cloth_anchor& a : c.anchors) {
        dist[a.particle] = 0.0f;
        from[a.particle] = a.particle;
        open.push({0.0f, a.particle});
    }
    while (!open.empty()) {
        const Entry e = open.top();
        open.pop();
        if (e.first > dist[e.second]) continue;
        for (uint32_t nb : adj[e.second]) {
            const float cand = e.first + length(sub(v(c.particles[nb].position), v(c.particles[e.second].position)));
            if (cand >= dist[nb]) continue;
            dist[nb] = cand;
            from[nb] = from[e.second];
            open.push({cand, nb});
        }
    }
    const float scale = std::max(1.0f, c.params.tether_scale);
    for (uint32_t i = 0; i < n; ++i) {
        if (c.particles[i].inv_mass == 0.0f || dist[i] >= inf) continue;
        c.tethers.push_back(pxg_cloth_tether{i, from[i], dist[i] * scale, 0u});
    }
}

// hair::StrandDesc (hair/strand_builder.h:29-37): world points root first, root orientation
// (x, y, z, w), particle mass
struct StrandDesc {
    std::vector<float> points;  // 3 per point
    float root_orientation[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    float particle_mass = 0.001f;
};

// make_strands (hair/strand_builder.h:58-128): roots pinned, each chain's segments as stretch
// constraints in two particle-disjoint batches (even segments of every strand, then odd), the
// rest pose in the root's frame and as offsets from the previous particle.
inline StrandSet make_strands(const std::vector<StrandDesc>& descs, const pxg_strand_params& params) {
    using namespace soft_math;
    StrandSet s;
    s.params = params;
    size_t total = 0;
    for (const StrandDesc& d : descs)
        if (d.points.size() >= 6) total += d.points.size() / 3;
    if (total == 0) return s;
    std::vector<pxg_distance_constraint> even, odd;
    for (const StrandDesc& d : descs) {
        const size_t np = d.points.size() / 3;
        if (np < 2) continue;
        pxg_strand st;
        std::memset(&st, 0, sizeof(st));
        st.first = static_cast<uint32_t>(s.particles.size());
        st.count = static_cast<uint32_t>(np);
        std::memcpy(st.root_position, &d.points[0], 3 * sizeof(float));
        std::memcpy(st.prev_root_position, &d.points[0], 3 * sizeof(float));
        std::memcpy(st.root_orientation, d.root_orientation, 4 * sizeof(float));
        std::memcpy(st.prev_root_orientation, d.root_orientation, 4 * sizeof(float));
        float to_root[4];
        inverse(d.root_orientation, to_root);
        const float im = d.particle_mass > 0.0f ? 1.0f / d.particle_mass : 1.0f;
        const V root = v(&d.points[0]);
        for (size_t k = 0; k < np; ++k) {
            const V pk = v(&d.points[3 * k]);
            pxg_particle p;
            std::memset(&p, 0, sizeof(p));
            put(p.position, pk);
            put(p.prev_position, pk);
            p.inv_mass = k == 0 ? 0.0f : im;
            s.particles.push_back(p);
            const V rl = rotate(to_root, sub(pk, root));
            s.rest_local.insert(s.rest_local.end(), {rl.x, rl.y, rl.z});
            const V ro = k == 0 ? V{0.0f, 0.0f, 0.0f} : sub(pk, v(&d.points[3 * (k - 1)]));
            s.rest_offset.insert(s.rest_offset.end(), {ro.x, ro.y, ro.z});
            if (k > 0) {
                pxg_distance_constraint c{st.first + uint32_t(k - 1), st.first + uint32_t(k),
                                          length(sub(pk, v(&d.points[3 * (k - 1)]))), 0.0f};
                ((k - 1) % 2 == 0 ? even : odd).push_back(c);
            }
        }
        s.strands.push_back(st);
    }
    s.stretch.insert(s.stretch.end(), even.begin(), even.end());
    s.stretch.insert(s.stretch.end(), odd.begin(), odd.end());
    const uint32_t ne = static_cast<uint32_t>(even.size()), no = static_cast<uint32_t>(odd.size());
    if (ne) s.stretch_batches.insert(s.stretch_batches.end(), {0u, ne});
    if (no) s.stretch_batches.insert(s.stretch_batches.end(), {ne, ne + no});
    compute_bounds(s.particles, s.bounds_min, s.bounds_max);
    return s;
}

// set_strand_roots (hair/strand_builder.h:139-150): positions 3 and orientations 4 (x, y, z, w)
// floats per strand, in strands order; orientations may be empty
inline void set_strand_roots(StrandSet& s, const std::vector<float>& positions, const std::vector<float>& orientations = {}) {
    const size_t n = std::min(s.strands.size(), positions.size() / 3);
    for (size_t i = 0; i < n; ++i) {
        pxg_strand& st = s.strands[i];
        std::memcpy(st.prev_root_position, st.root_position, sizeof(st.root_position));
        std::memcpy(st.prev_root_orientation, st.root_orientation, sizeof(st.root_orientation));
        std::memcpy(st.root_position, &positions[3 * i], 3 * sizeof(float));
        if (4 * i + 3 < orientations.size()) std::memcpy(st.root_orientation, &orientations[4 * i], 4 * sizeof(float));
    }
}

}  // namespace pxg
