// C++ facade over the pxg C ABI with the reference's PhysicsWorld interface
// (physxcoopa/world.h). A caller of physxcoopa::PhysicsWorld switches by including this header
// and linking libpxg.so; bodies and shapes are the ABI records of pxg.h.
//
// State model: the GPU owns body state between steps. The host keeps a mirror that is pulled
// on demand (get_body, body_count, ...) and pushed before the next step when it was handed out
// for writing or the body set changed; a push uploads only the bodies whose records changed.
// get_body() and pose() may be called from several threads at once between steps (the first
// call after a step downloads, the others wait for it), as the reference's job-parallel
// transform write-back does (system/physics_system.h:1058-1075). Errors follow the reference's conventions: stale handles
// are no-ops or nullptr (world.h:340-364); a failure inside the library throws
// std::runtime_error("[physxcoopa] ...") (util/error.h:24-26).
#pragma once

#include "pxg.h"
#include "pxg_soft.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace pxg {

// Generational handle (dynamics/body.h:61-77).
struct BodyId {
    static constexpr uint32_t kInvalid = 0xFFFFFFFFu;
    uint32_t index = kInvalid;
    uint32_t generation = 0;
    bool is_valid() const { return index != kInvalid; }
    bool operator==(const BodyId& o) const { return index == o.index && generation == o.generation; }
    bool operator!=(const BodyId& o) const { return !(*this == o); }
};

// 32 x 32 symmetric collision matrix, all pairs enabled by default (broadphase/layer_matrix.h:21-47).
class LayerMatrix {
public:
    LayerMatrix() {
        for (auto& m : mask_) m = 0xFFFFFFFFu;
    }
    bool should_collide(uint32_t a, uint32_t b) const { return (mask_[a & 31u] >> (b & 31u)) & 1u; }
    void set_layer_collision(uint32_t a, uint32_t b, bool collide) {
        a &= 31u;
        b &= 31u;
        if (collide) {
            mask_[a] |= 1u << b;
            mask_[b] |= 1u << a;
        } else {
            mask_[a] &= ~(1u << b);
            mask_[b] &= ~(1u << a);
        }
        ++version_;
    }
    const uint32_t* rows() const { return mask_; }
    uint64_t version() const { return version_; }

private:
    uint32_t mask_[32];
    uint64_t version_ = 0;
};

// The slot list the reference's on_substep uses (coopa::event::Signal<PhysicsWorld&, float>).
template <class... Args>
class Signal {
public:
    using Slot = std::function<void(Args...)>;
    uint64_t connect(Slot s) {
        slots_.emplace_back(++next_, std::move(s));
        return next_;
    }
    void disconnect(uint64_t id) {
        for (size_t i = 0; i < slots_.size(); ++i)
            if (slots_[i].first == id) {
                slots_.erase(slots_.begin() + static_cast<std::ptrdiff_t>(i));
                return;
            }
    }
    bool empty() const { return slots_.empty(); }
    void emit(Args... a) {
        for (auto& s : slots_) s.second(a...);
    }

private:
    std::vector<std::pair<uint64_t, Slot>> slots_;
    uint64_t next_ = 0;
};

[[noreturn]] inline void raise_error(const std::string& what) {
    throw std::runtime_error("[physxcoopa] " + what + ": " + pxg_last_error());
}
inline void check(int status, const char* what) {
    if (status != PXG_OK) raise_error(what);
}

inline pxg_config default_config() {
    pxg_config c;
    pxg_config_default(&c);
    return c;
}
// A disabled shape: the body takes part in dynamics but never in collision (world.h:85-95).
// Generational joint handle (dynamics/joint.h:39-52).
struct JointId {
    static constexpr uint32_t kInvalid = 0xFFFFFFFFu;
    uint32_t index = kInvalid;
    uint32_t generation = 0;
    bool is_valid() const { return index != kInvalid; }
    bool operator==(const JointId& o) const { return index == o.index && generation == o.generation; }
    bool operator!=(const JointId& o) const { return !(*this == o); }
};

struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};

// geometry::Ray (geometry/ray.h:32-35), query::RaycastHit and query::QueryFilter without its
// predicate (query/queries.h:56-98)
struct Ray {
    Vec3 origin;
    Vec3 direction{0.0f, 0.0f, -1.0f};  // unit length
    float max_distance = 3.402823466e+38f;
};
struct Quat {
    float x = 0.0f, y = 0.0f, z = 0.0f, w = 1.0f;
};
struct RaycastHit {
    Vec3 point;
    Vec3 normal{0.0f, 0.0f, 1.0f};
    float distance = 0.0f;
    BodyId body;
    uint32_t shape_index = 0xFFFFFFFFu;
    bool started_inside = false;  // shape casts only
};
// One overlap of compute_penetration() (query::Penetration, query/queries.h:111-117): moving the
// query shape by normal * depth separates it; normal points from the target toward the query shape.
struct Penetration {
    Vec3 normal{0.0f, 0.0f, 1.0f};
    float depth = 0.0f;
    Vec3 point;  // a representative contact point on the target
    BodyId body;
    uint32_t shape_index = 0xFFFFFFFFu;
};
struct QueryFilter {
    uint32_t layer_mask = ~0u;
    bool include_triggers = true;
    BodyId ignore;
};

// One pair's enter / stay / exit transition of one step (collision/contact_event.h:29-36).
struct ContactEvent {
    enum class Type { Enter, Stay, Exit };
    BodyId a;
    BodyId b;
    Type type = Type::Enter;
    bool is_trigger = false;
};

// glm quaternion arithmetic in float, glm's term order (x, y, z, w storage) -- for the joints'
// rest rotation, inverse(qa) * qb at creation (world.h:415).
namespace quat_ops {
inline void mul(const float* p, const float* q, float* r) {
    const float w = p[3] * q[3] - p[0] * q[0] - p[1] * q[1] - p[2] * q[2];
    const float x = p[3] * q[0] + p[0] * q[3] + p[1] * q[2] - p[2] * q[1];
    const float y = p[3] * q[1] + p[1] * q[3] + p[2] * q[0] - p[0] * q[2];
    const float z = p[3] * q[2] + p[2] * q[3] + p[0] * q[1] - p[1] * q[0];
    r[0] = x; r[1] = y; r[2] = z; r[3] = w;
}
inline void inverse(const float* q, float* r) {
    const float d = (q[0] * q[0] + q[1] * q[1]) + (q[2] * q[2] + q[3] * q[3]);
    r[0] = -q[0] / d; r[1] = -q[1] / d; r[2] = -q[2] / d; r[3] = q[3] / d;
}
inline void rotate(const float* q, const float* v, float* r) {
    auto cross = [](const float* a, const float* b, float* c) {
        c[0] = a[1] * b[2] - b[1] * a[2];
        c[1] = a[2] * b[0] - b[2] * a[0];
        c[2] = a[0] * b[1] - b[0] * a[1];
    };
    float uv[3], uuv[3];
    cross(q, v, uv);
    cross(q, uv, uuv);
    for (int i = 0; i < 3; ++i) r[i] = v[i] + ((uv[i] * q[3]) + uuv[i]) * 2.0f;
}
}  // namespace quat_ops

inline pxg_shape no_shape() {
    pxg_shape s;
    std::memset(&s, 0, sizeof(s));
    s.local_rotation[3] = 1.0f;
    return s;
}

class PhysicsWorld {
public:
    // Fired once per step_fixed() after force integration and contact generation, before the
    // solve (world.h:70-78, 1299). Slots may change velocities and forces through get_body().
    Signal<PhysicsWorld&, float> on_substep;

    explicit PhysicsWorld(const pxg_config& config = default_config()) : cfg_(config) {
        // device buffers are sized once per handle: reserve room to add bodies later
        if (cfg_.max_bodies == 0) cfg_.max_bodies = 256;
        if (cfg_.max_shapes == 0) cfg_.max_shapes = cfg_.max_bodies;
        check(pxg_create(&cfg_, &w_), "pxg_create");
        applied_cfg_ = cfg_;
    }
    ~PhysicsWorld() {
        if (w_) pxg_destroy(w_);
    }
    PhysicsWorld(const PhysicsWorld&) = delete;
    PhysicsWorld& operator=(const PhysicsWorld&) = delete;

    // --- body lifecycle (world.h:95-143, 340-364) ---
    // From an on_substep slot the insertion is deferred to the top of the next step_fixed() and
    // the returned handle is invalid (world.h:281-300).
    BodyId add_body(const pxg_body& initial, const pxg_shape& shape = no_shape()) {
        if (in_substep_) {
            deferred_.push_back({true, BodyId{}, initial, shape});
            return BodyId{};
        }
        pull();
        uint32_t index;
        if (!free_.empty()) {
            index = free_.back();
            free_.pop_back();
        } else {
            index = static_cast<uint32_t>(bodies_.size());
            bodies_.push_back(pxg_body{});
            shapes_.push_back(no_shape());
            generations_.push_back(0);
        }
        pxg_body& b = bodies_[index];
        b = initial;
        b.alive = 1;
        b.generation = generations_[index];  // cloth anchors and groom collide lists name bodies by BodyId
        std::memcpy(b.prev_position, b.position, sizeof(b.position));
        std::memcpy(b.prev_orientation, b.orientation, sizeof(b.orientation));
        pxg_shape s = shape;
        s.owner = index;
        shapes_[index] = s;
        new_shapes_.push_back(index);
        structure_dirty_ = true;
        return BodyId{index, generations_[index]};
    }

    void remove_body(BodyId id) {
        if (!is_valid(id)) return;
        // joints on the body stop acting at once, and its ignored pairs are forgotten, also when
        // the removal itself is deferred (world.h:130-142, 331-333)
        pull_joints();
        for (auto& j : joints_) {
            if (j.valid && (j.body_a == id.index || j.body_b == id.index)) {
                j.valid = 0;
                joints_dirty_ = true;
            }
        }
        for (size_t k = 0; k < ignored_.size();) {
            if (ignored_[k].first == id.index || ignored_[k].second == id.index) {
                ignored_.erase(ignored_.begin() + long(k));
                ignored_dirty_ = true;
            } else {
                ++k;
            }
        }
        if (in_substep_) {
            deferred_.push_back({false, id, pxg_body{}, no_shape()});
            return;
        }
        pull();
        bodies_[id.index].alive = 0;
        shapes_[id.index].enabled = 0;
        ++generations_[id.index];
        free_.push_back(id.index);
        structure_dirty_ = true;
    }

    bool is_valid(BodyId id) const {
        return id.is_valid() && id.index < bodies_.size() && generations_[id.index] == id.generation &&
               alive_on_host(id.index);
    }

    // The body `id` addresses, or nullptr for a stale handle (world.h:347). Writes through the
    // pointer reach the GPU before the next step.
    pxg_body* get_body(BodyId id) {
        if (!is_valid(id)) return nullptr;
        pull();
        bodies_dirty_.store(true, std::memory_order_relaxed);
        return &bodies_[id.index];
    }
    const pxg_body* get_body(BodyId id) const {
        if (!is_valid(id)) return nullptr;
        const_cast<PhysicsWorld*>(this)->pull();
        return &bodies_[id.index];
    }
    // What a transform write-back reads of `id` (type, interpolation mode, previous and current
    // pose), or nullptr for a stale handle: the poses of all bodies are downloaded once per step,
    // 64 bytes per body, by the first call after the step (thread-safe like get_body).
    const pxg_pose* pose(BodyId id) const {
        if (!is_valid(id)) return nullptr;
        const_cast<PhysicsWorld*>(this)->pull_poses();
        return &poses_[id.index];
    }
    // Fractional progress through the next substep after the last step(dt), in [0, 1), for
    // render interpolation (world.h:1240, 1319-1320).
    float interpolation_alpha() const { return interpolation_alpha_; }
    // Body records the last step_fixed() uploaded before stepping (0 when nothing changed on the
    // host; the full count after a structural change).
    uint32_t bodies_uploaded_last_step() const { return bodies_uploaded_; }

    // The primary shape of `id` (world.h:367-374); edits apply before the next step.
    pxg_shape* get_shape(BodyId id) {
        if (!is_valid(id)) return nullptr;
        structure_dirty_ = true;
        new_shapes_.push_back(id.index);  // its proxy is re-created
        return &shapes_[id.index];
    }

    // --- joints (world.h:378-518) ---
    // A hinge: b's local_axis_b stays aligned with a's local_axis_a; with use_limits the twist
    // about it stays in [min_angle, max_angle] from the relative orientation at creation.
    JointId add_hinge_joint(BodyId a, BodyId b, Vec3 local_anchor_a, Vec3 local_anchor_b, Vec3 local_axis_a,
                            Vec3 local_axis_b, bool use_limits = false, float min_angle = 0.0f, float max_angle = 0.0f,
                            bool collide_connected = false) {
        if (!is_valid(a) || !is_valid(b)) return JointId{};
        pxg_joint j = joint_base(PXG_HINGE, a, b, local_anchor_a, local_anchor_b, collide_connected);
        set3(j.local_axis_a, local_axis_a);
        set3(j.local_axis_b, local_axis_b);
        j.use_limits = use_limits ? 1u : 0u;
        j.min_angle = min_angle;
        j.max_angle = max_angle;
        return insert_joint(j);
    }
    JointId add_ball_joint(BodyId a, BodyId b, Vec3 local_anchor_a, Vec3 local_anchor_b, bool collide_connected = false) {
        if (!is_valid(a) || !is_valid(b)) return JointId{};
        return insert_joint(joint_base(PXG_BALL, a, b, local_anchor_a, local_anchor_b, collide_connected));
    }
    // b's twist axis is a's carried through the rest pose (world.h:456-477).
    JointId add_cone_twist_joint(BodyId a, BodyId b, Vec3 local_anchor_a, Vec3 local_anchor_b, Vec3 local_twist_axis_a,
                                 float swing_limit, float twist_min, float twist_max, bool collide_connected = false) {
        if (!is_valid(a) || !is_valid(b)) return JointId{};
        pxg_joint j = joint_base(PXG_CONE_TWIST, a, b, local_anchor_a, local_anchor_b, collide_connected);
        float ax[3] = {local_twist_axis_a.x, local_twist_axis_a.y, local_twist_axis_a.z};
        const float d = (ax[0] * ax[0] + ax[1] * ax[1]) + ax[2] * ax[2];
        const float inv = 1.0f / std::sqrt(d);
        for (float& c : ax) c = c * inv;
        std::memcpy(j.local_axis_a, ax, sizeof(ax));
        float rinv[4];
        quat_ops::inverse(j.rest_relative_rotation, rinv);
        quat_ops::rotate(rinv, ax, j.local_axis_b);
        j.swing_limit = swing_limit;
        j.twist_min = twist_min;
        j.twist_max = twist_max;
        return insert_joint(j);
    }
    // Inserts a fully described joint; its accumulators start at zero (world.h:485-493).
    JointId add_joint(const pxg_joint& desc) {
        if (!is_valid(BodyId{desc.body_a, generations_at(desc.body_a)}) ||
            !is_valid(BodyId{desc.body_b, generations_at(desc.body_b)}))
            return JointId{};
        pxg_joint j = desc;
        std::memset(j.point_impulse, 0, sizeof(j.point_impulse));
        std::memset(j.axis_impulse, 0, sizeof(j.axis_impulse));
        j.limit_impulse = j.swing_impulse = 0.0f;
        j.valid = 1;
        return insert_joint(j);
    }
    bool is_valid(JointId id) const {
        return id.is_valid() && id.index < joints_.size() && joint_alive_[id.index] &&
               joint_generations_[id.index] == id.generation;
    }
    void remove_joint(JointId id) {  // world.h:502-510
        if (!is_valid(id)) return;
        pull_joints();
        joints_[id.index].valid = 0;
        joint_alive_[id.index] = 0;
        ++joint_generations_[id.index];
        free_joints_.push_back(id.index);
        joints_dirty_ = true;
    }
    // The joint (with the solver's accumulators), or nullptr for a stale handle (world.h:514).
    // Writes through the pointer reach the GPU before the next step.
    pxg_joint* get_joint(JointId id) {
        if (!is_valid(id)) return nullptr;
        pull_joints();
        joints_dirty_ = true;
        return &joints_[id.index];
    }
    const pxg_joint* get_joint(JointId id) const {
        if (!is_valid(id)) return nullptr;
        const_cast<PhysicsWorld*>(this)->pull_joints();
        return &joints_[id.index];
    }
    // Contacts between two bodies off (collide false) or back on, regardless of layers
    // (world.h:527-536).
    void set_pair_collision(BodyId a, BodyId b, bool collide) {
        if (!is_valid(a) || !is_valid(b) || a == b) return;
        const std::pair<uint32_t, uint32_t> key(std::min(a.index, b.index), std::max(a.index, b.index));
        for (size_t k = 0; k < ignored_.size(); ++k) {
            if (ignored_[k] == key) {
                if (collide) {
                    ignored_.erase(ignored_.begin() + long(k));
                    ignored_dirty_ = true;
                }
                return;
            }
        }
        if (!collide) {
            ignored_.push_back(key);
            ignored_dirty_ = true;
        }
    }
    bool pair_collides(BodyId a, BodyId b) const {
        const std::pair<uint32_t, uint32_t> key(std::min(a.index, b.index), std::max(a.index, b.index));
        for (const auto& p : ignored_)
            if (p == key) return false;
        return true;
    }

    uint32_t body_count() const {
        uint32_t n = 0;
        for (size_t i = 0; i < bodies_.size(); ++i) n += alive_on_host(static_cast<uint32_t>(i)) ? 1u : 0u;
        return n;
    }

    LayerMatrix& layers() { return layers_; }
    const LayerMatrix& layers() const { return layers_; }

    // Mutable like the reference's config() (world.h:1207); changes apply at the next step.
    pxg_config& config() { return cfg_; }
    const pxg_config& config() const { return cfg_; }

    void set_gravity(float x, float y, float z) {
        cfg_.gravity[0] = x;
        cfg_.gravity[1] = y;
        cfg_.gravity[2] = z;
    }
    const float* gravity() const { return cfg_.gravity; }

    // --- stepping (world.h:1229-1317) ---
    // Accumulates dt (clamped to 0.25 s) and runs up to max_substeps fixed substeps.
    uint32_t step(float dt) {
        clear_events();  // world.h:1230
        const float frame = dt < 0.25f ? dt : 0.25f;
        accumulator_ += frame;
        const float h = cfg_.fixed_dt;
        uint32_t n = 0;
        while (accumulator_ >= h && n < cfg_.max_substeps) {
            step_fixed(h);
            accumulator_ -= h;
            ++n;
        }
        interpolation_alpha_ = accumulator_ / h;
        if (frame > 0.0f && (live_count(cloth_alive_) || live_count(strand_alive_))) {  // world.h:1241-1256
            push();
            check(pxg_step_soft(w_, frame, n), "step (cloth and hair)");
            soft_on_device_ = true;
        }
        return n;
    }

    void step_fixed(float h) {
        std::vector<Deferred> cmds;
        cmds.swap(deferred_);
        for (const Deferred& c : cmds) {
            if (c.create) add_body(c.body, c.shape);
            else remove_body(c.id);
        }
        push();
        if (on_substep.empty()) {
            check(pxg_step_fixed(w_, h), "step_fixed");
        } else {
            check(pxg_step_phase(w_, PXG_PHASE_FORCES_COLLIDE, h), "step_fixed (collide)");
            state_on_device_.store(true, std::memory_order_release);
            poses_valid_.store(false, std::memory_order_release);
            bodies_dirty_.store(false, std::memory_order_relaxed);
            in_substep_ = true;
            on_substep.emit(*this, h);
            in_substep_ = false;
            if (bodies_dirty_.load(std::memory_order_relaxed)) {
                upload_changed_bodies("on_substep upload");
                bodies_dirty_.store(false, std::memory_order_relaxed);
            }
            push_joints();
            check(pxg_step_phase(w_, PXG_PHASE_SOLVE_INTEGRATE, h), "step_fixed (solve)");
        }
        state_on_device_.store(true, std::memory_order_release);
        poses_valid_.store(false, std::memory_order_release);
        joints_on_device_ = !joints_.empty();
        manifolds_valid_ = false;
        events_valid_ = false;
    }

    // Contact events queued since the last clear_events() (world.h:744): every step_fixed()
    // appends its enters, stays and exits; step() clears the queue first. A capacity growth
    // (more bodies than the handle holds) starts a new handle, whose first step reports its
    // contacts as enters.
    const std::vector<ContactEvent>& events() const {
        if (!events_valid_) {
            uint32_t n = 0;
            check(pxg_get_events(w_, nullptr, 0, &n), "events");
            std::vector<pxg_contact_event> raw(n);
            if (n) check(pxg_get_events(w_, raw.data(), n, &n), "events");
            events_.clear();
            for (uint32_t i = 0; i < n && i < raw.size(); ++i) {
                ContactEvent e;
                e.a = BodyId{raw[i].a, raw[i].a < generations_.size() ? generations_[raw[i].a] : 0u};
                e.b = BodyId{raw[i].b, raw[i].b < generations_.size() ? generations_[raw[i].b] : 0u};
                e.type = raw[i].type == PXG_EVENT_ENTER ? ContactEvent::Type::Enter
                         : raw[i].type == PXG_EVENT_STAY ? ContactEvent::Type::Stay
                                                         : ContactEvent::Type::Exit;
                e.is_trigger = raw[i].is_trigger != 0;
                events_.push_back(e);
            }
            events_valid_ = true;
        }
        return events_;
    }
    // --- scene queries (world.h:766-847), against the bodies' current poses ---
    // Closest hit along the ray (equal distances: the lower shape slot).
    bool raycast(const Ray& ray, RaycastHit& hit, const QueryFilter& filter = QueryFilter()) {
        std::vector<Ray> one(1, ray);
        std::vector<RaycastHit> hits;
        std::vector<uint8_t> found;
        raycast_batch(one, hits, found, filter);
        if (!found[0]) return false;
        hit = hits[0];
        return true;
    }
    bool raycast(const Ray& ray, RaycastHit& hit, uint32_t layer_mask, bool include_triggers = true) {
        QueryFilter f;
        f.layer_mask = layer_mask;
        f.include_triggers = include_triggers;
        return raycast(ray, hit, f);
    }
    bool raycast_any(const Ray& ray, const QueryFilter& filter = QueryFilter()) {
        push();
        pxg_ray r = to_ray(ray);
        pxg_query_filter f = to_filter(filter);
        uint32_t any = 0;
        check(pxg_raycast_any(w_, &r, 1, &f, &any), "raycast_any");
        return any != 0;
    }
    // Every hit along the ray, nearest first.
    std::vector<RaycastHit> raycast_all(const Ray& ray, const QueryFilter& filter = QueryFilter()) {
        push();
        pxg_ray r = to_ray(ray);
        pxg_query_filter f = to_filter(filter);
        uint32_t n = 0;
        check(pxg_raycast_all(w_, &r, &f, nullptr, 0, &n), "raycast_all");
        std::vector<pxg_raycast_hit> raw(n);
        if (n) check(pxg_raycast_all(w_, &r, &f, raw.data(), n, &n), "raycast_all");
        std::vector<RaycastHit> out;
        for (uint32_t i = 0; i < n && i < raw.size(); ++i) out.push_back(from_hit(raw[i]));
        return out;
    }
    // Bodies whose shape overlaps a world-space sphere / OBB / capsule (world.h:1002-1060), once
    // per overlapping shape, in shape-slot order.
    std::vector<BodyId> overlap_sphere(Vec3 center, float radius, const QueryFilter& filter = QueryFilter()) {
        pxg_overlap_query q = overlap_base(PXG_SPHERE);
        set3(q.center, center);
        q.radius = radius;
        return overlap(q, filter);
    }
    std::vector<BodyId> overlap_box(Vec3 center, Vec3 half_extents, Quat orientation,
                                    const QueryFilter& filter = QueryFilter()) {
        pxg_overlap_query q = overlap_base(PXG_BOX);
        set3(q.center, center);
        set3(q.half_extents, half_extents);
        q.orientation[0] = orientation.x;
        q.orientation[1] = orientation.y;
        q.orientation[2] = orientation.z;
        q.orientation[3] = orientation.w;
        return overlap(q, filter);
    }
    std::vector<BodyId> overlap_capsule(Vec3 a, Vec3 b, float radius, const QueryFilter& filter = QueryFilter()) {
        pxg_overlap_query q = overlap_base(PXG_CAPSULE);
        set3(q.a, a);
        set3(q.b, b);
        q.radius = radius;
        return overlap(q, filter);
    }

    // The first body a swept sphere / box / capsule touches (sphere_cast(), box_cast(),
    // capsule_cast(), world.h:911-950).
    bool sphere_cast(Vec3 origin, float radius, Vec3 dir, float max_distance, RaycastHit& hit,
                     const QueryFilter& filter = QueryFilter()) {
        pxg_cast_query c = cast_base(max_distance, origin, Quat{}, dir);
        c.shape.type = PXG_SPHERE;
        c.shape.radius = radius;
        return shape_cast(c, hit, filter);
    }
    bool box_cast(Vec3 origin, Vec3 half_extents, Quat orientation, Vec3 dir, float max_distance, RaycastHit& hit,
                  const QueryFilter& filter = QueryFilter()) {
        pxg_cast_query c = cast_base(max_distance, origin, orientation, dir);
        c.shape.type = PXG_BOX;
        set3(c.shape.half_extents, half_extents);
        return shape_cast(c, hit, filter);
    }
    bool capsule_cast(Vec3 origin, float radius, float half_height, int direction_axis, Quat orientation, Vec3 dir,
                      float max_distance, RaycastHit& hit, const QueryFilter& filter = QueryFilter()) {
        pxg_cast_query c = cast_base(max_distance, origin, orientation, dir);
        c.shape.type = PXG_CAPSULE;
        c.shape.capsule_radius = radius;
        c.shape.capsule_half_height = half_height;
        c.shape.capsule_axis = direction_axis;
        return shape_cast(c, hit, filter);
    }

    // Every shape that a sphere, box or capsule posed at pos / rot penetrates, with its push-out, in
    // shape-slot order (compute_penetration(), world.h:963-995). shape.type must be PXG_SPHERE,
    // PXG_BOX or PXG_CAPSULE; its owner, material and layer are not used.
    std::vector<Penetration> compute_penetration(const pxg_shape& shape, Vec3 pos, Quat rot,
                                                 const QueryFilter& filter = QueryFilter()) {
        push();
        pxg_query_filter f = to_filter(filter);
        const float p[3] = {pos.x, pos.y, pos.z}, q[4] = {rot.x, rot.y, rot.z, rot.w};
        uint32_t n = 0;
        check(pxg_compute_penetration(w_, &shape, p, q, &f, nullptr, 0, &n), "compute_penetration");
        std::vector<pxg_penetration> raw(n);
        if (n) check(pxg_compute_penetration(w_, &shape, p, q, &f, raw.data(), n, &n), "compute_penetration");
        std::vector<Penetration> out;
        for (uint32_t i = 0; i < n && i < raw.size(); ++i) {
            Penetration e;
            e.normal = Vec3{raw[i].normal[0], raw[i].normal[1], raw[i].normal[2]};
            e.depth = raw[i].depth;
            e.point = Vec3{raw[i].point[0], raw[i].point[1], raw[i].point[2]};
            const uint32_t b = raw[i].body;
            e.body = BodyId{b, b < generations_.size() ? generations_[b] : 0u};
            e.shape_index = raw[i].shape_index;
            out.push_back(e);
        }
        return out;
    }

    // Many overlap queries in one GPU pass (no counterpart in the reference, which answers one per
    // call): out[i] lists what the single query returns (all of it when it fits in cap entries;
    // counts[i] is always the full count).
    void overlap_batch(const std::vector<pxg_overlap_query>& queries, std::vector<std::vector<BodyId>>& out,
                       std::vector<uint32_t>& counts, uint32_t cap = 64, const QueryFilter& filter = QueryFilter()) {
        push();
        pxg_query_filter f = to_filter(filter);
        const uint32_t n = static_cast<uint32_t>(queries.size());
        std::vector<uint32_t> raw(size_t(n) * cap);
        counts.assign(n, 0u);
        if (n) check(pxg_overlap_batch(w_, queries.data(), n, &f, raw.data(), cap, counts.data()), "overlap_batch");
        out.assign(n, std::vector<BodyId>());
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t k = 0; k < counts[i] && k < cap; ++k) {
                const uint32_t b = raw[size_t(i) * cap + k];
                out[i].push_back(BodyId{b, b < generations_.size() ? generations_[b] : 0u});
            }
    }

    // Many shape casts in one GPU pass (no counterpart in the reference, which casts one shape per
    // call): hits[i] is valid where found[i] != 0 and equals what the single cast returns.
    void shape_cast_batch(const std::vector<pxg_cast_query>& casts, std::vector<RaycastHit>& hits,
                          std::vector<uint8_t>& found, const QueryFilter& filter = QueryFilter()) {
        push();
        pxg_query_filter f = to_filter(filter);
        std::vector<pxg_raycast_hit> raw(casts.size());
        if (!casts.empty())
            check(pxg_shape_cast_batch(w_, casts.data(), static_cast<uint32_t>(casts.size()), &f, raw.data()),
                  "shape_cast_batch");
        hits.assign(casts.size(), RaycastHit());
        found.assign(casts.size(), 0);
        for (size_t i = 0; i < raw.size(); ++i) {
            if (!raw[i].hit) continue;
            hits[i] = from_hit(raw[i]);
            hits[i].started_inside = raw[i].started_inside != 0;
            found[i] = 1;
        }
    }

    // The closest hits of many rays in one GPU pass (no counterpart in the reference, which casts
    // one ray per call): hits[i] is valid where found[i] != 0.
    void raycast_batch(const std::vector<Ray>& rays, std::vector<RaycastHit>& hits, std::vector<uint8_t>& found,
                       const QueryFilter& filter = QueryFilter()) {
        push();
        std::vector<pxg_ray> r(rays.size());
        for (size_t i = 0; i < rays.size(); ++i) r[i] = to_ray(rays[i]);
        std::vector<pxg_raycast_hit> raw(rays.size());
        pxg_query_filter f = to_filter(filter);
        check(pxg_raycast(w_, r.data(), static_cast<uint32_t>(r.size()), &f, raw.data()), "raycast");
        hits.assign(rays.size(), RaycastHit());
        found.assign(rays.size(), 0);
        for (size_t i = 0; i < raw.size(); ++i) {
            if (!raw[i].hit) continue;
            hits[i] = from_hit(raw[i]);
            found[i] = 1;
        }
    }

    void clear_events() {  // world.h:748
        check(pxg_clear_events(w_), "clear_events");
        events_.clear();
        events_valid_ = true;
    }

    // This step's valid manifolds with the solver's final impulses (world.h:731).
    const std::vector<pxg_manifold>& manifolds() const {
        if (!manifolds_valid_) {
            uint32_t n = 0;
            check(pxg_get_manifolds(w_, nullptr, 0, &n), "manifolds");
            manifolds_.resize(n);
            if (n) check(pxg_get_manifolds(w_, manifolds_.data(), n, &n), "manifolds");
            manifolds_.resize(n);
            manifolds_valid_ = true;
        }
        return manifolds_;
    }

    // FNV-1a over alive bodies' pose and velocities (world.h:1334-1359).
    uint64_t world_state_hash() {
        push();
        uint64_t h = 0;
        check(pxg_state_hash(w_, &h), "world_state_hash");
        return h;
    }

    pxg_stats stats() const {
        pxg_stats s;
        check(pxg_get_stats(w_, &s), "stats");
        return s;
    }

    // The convex hulls and triangle meshes that shapes index (Shape::hull / Shape::mesh,
    // collision/shape.h:84-98): PXG_CONVEX_HULL / PXG_TRIANGLE_MESH shapes name a table entry in
    // capsule_axis and their scale in radius. The arrays are copied; the tables go up before the
    // next shape upload (and again whenever the world is re-created to grow).
    void set_hulls(const pxg_hull* hulls, uint32_t n) {
        hull_arrays_.assign(n, HullArrays{});
        for (uint32_t k = 0; k < n; ++k) {
            const pxg_hull& h = hulls[k];
            HullArrays& a = hull_arrays_[k];
            a.desc = h;
            a.v.assign(h.vertices, h.vertices + 3 * size_t(h.n_vertices));
            a.planes.assign(h.face_planes, h.face_planes + 4 * size_t(h.n_faces));
            a.starts.assign(h.face_starts, h.face_starts + h.n_faces + 1);
            a.fv.assign(h.face_vertices, h.face_vertices + (h.n_faces ? h.face_starts[h.n_faces] : 0u));
            if (h.n_edges) a.edges.assign(h.edges, h.edges + 2 * size_t(h.n_edges));
        }
        geometry_dirty_ = structure_dirty_ = true;
    }
    void set_meshes(const pxg_mesh_desc* meshes, uint32_t n) {
        mesh_arrays_.assign(n, MeshArrays{});
        for (uint32_t k = 0; k < n; ++k) {
            mesh_arrays_[k].positions.assign(meshes[k].positions, meshes[k].positions + 3 * size_t(meshes[k].n_positions));
            mesh_arrays_[k].indices.assign(meshes[k].indices, meshes[k].indices + meshes[k].n_indices);
        }
        geometry_dirty_ = structure_dirty_ = true;
    }

    // --- cloth and hair (world.h:552-700) ---
    // The tables live on the GPU between frames: the first get_cloth / get_strands /
    // for_each_* after a step downloads the alive entries; handing them out for writing sends the
    // tables up again before the next step. Anchors and collide lists name bodies by BodyId
    // (index, generation); a stale one pins in place / is skipped, as in the reference.
    ClothId add_cloth(Cloth&& c) {
        if (c.particles.empty()) return ClothId{};
        pull_soft();
        compute_bounds(c.particles, c.bounds_min, c.bounds_max);
        const uint32_t index = take_slot(cloths_, cloth_generations_, cloth_alive_, free_cloths_);
        cloths_[index] = std::move(c);
        cloths_dirty_ = true;
        return ClothId{index, cloth_generations_[index]};
    }
    bool is_valid(ClothId id) const {
        return id.is_valid() && id.index < cloths_.size() && cloth_alive_[id.index] &&
               cloth_generations_[id.index] == id.generation;
    }
    void remove_cloth(ClothId id) {
        if (!is_valid(id)) return;
        pull_soft();
        cloths_[id.index] = Cloth{};
        cloth_alive_[id.index] = 0;
        ++cloth_generations_[id.index];
        free_cloths_.push_back(id.index);
        cloths_dirty_ = true;
    }
    Cloth* get_cloth(ClothId id) {
        if (!is_valid(id)) return nullptr;
        pull_soft();
        cloths_dirty_ = true;
        return &cloths_[id.index];
    }
    const Cloth* get_cloth(ClothId id) const {
        if (!is_valid(id)) return nullptr;
        const_cast<PhysicsWorld*>(this)->pull_soft();
        return &cloths_[id.index];
    }
    std::size_t cloth_count() const { return live_count(cloth_alive_); }
    template <typename Fn>
    void for_each_cloth(Fn&& fn) {
        pull_soft();
        cloths_dirty_ = true;
        for (uint32_t i = 0; i < cloths_.size(); ++i)
            if (cloth_alive_[i]) fn(ClothId{i, cloth_generations_[i]}, cloths_[i]);
    }

    StrandSetId add_strands(StrandSet&& set) {
        if (set.particles.empty() || set.strands.empty()) return StrandSetId{};
        pull_soft();
        compute_bounds(set.particles, set.bounds_min, set.bounds_max);
        const uint32_t index = take_slot(strand_sets_, strand_generations_, strand_alive_, free_strands_);
        strand_sets_[index] = std::move(set);
        strands_dirty_ = true;
        return StrandSetId{index, strand_generations_[index]};
    }
    bool is_valid(StrandSetId id) const {
        return id.is_valid() && id.index < strand_sets_.size() && strand_alive_[id.index] &&
               strand_generations_[id.index] == id.generation;
    }
    void remove_strands(StrandSetId id) {
        if (!is_valid(id)) return;
        pull_soft();
        strand_sets_[id.index] = StrandSet{};
        strand_alive_[id.index] = 0;
        ++strand_generations_[id.index];
        free_strands_.push_back(id.index);
        strands_dirty_ = true;
    }
    StrandSet* get_strands(StrandSetId id) {
        if (!is_valid(id)) return nullptr;
        pull_soft();
        strands_dirty_ = true;
        return &strand_sets_[id.index];
    }
    const StrandSet* get_strands(StrandSetId id) const {
        if (!is_valid(id)) return nullptr;
        const_cast<PhysicsWorld*>(this)->pull_soft();
        return &strand_sets_[id.index];
    }
    std::size_t strand_set_count() const { return live_count(strand_alive_); }
    template <typename Fn>
    void for_each_strands(Fn&& fn) {
        pull_soft();
        strands_dirty_ = true;
        for (uint32_t i = 0; i < strand_sets_.size(); ++i)
            if (strand_alive_[i]) fn(StrandSetId{i, strand_generations_[i]}, strand_sets_[i]);
    }
    // A body's BodyId as a groom's collide-list entry or a cloth anchor's (index, generation).
    std::pair<uint32_t, uint32_t> body_ref(BodyId id) const { return {id.index, id.generation}; }

    pxg_world* handle() { return w_; }

private:
    // A new handle with room for at least n bodies; the warm-start cache moves over (the caller
    // restores the fat AABBs).
    void grow(uint32_t n) {
        pull_soft();  // the cloth and hair state on the old handle
        uint32_t cap = cfg_.max_bodies;
        while (cap < n) cap *= 2;
        uint32_t nk = 0;
        check(pxg_get_warm_cache(w_, nullptr, nullptr, 0, &nk), "warm cache");
        std::vector<uint64_t> keys(nk);
        std::vector<float> imp(size_t(nk) * 3);
        if (nk) check(pxg_get_warm_cache(w_, keys.data(), imp.data(), nk, &nk), "warm cache");
        pxg_config c = cfg_;
        c.max_bodies = c.max_shapes = cap;
        pxg_world* nw = nullptr;
        check(pxg_create(&c, &nw), "pxg_create");
        pxg_destroy(w_);
        w_ = nw;
        cfg_ = c;
        applied_cfg_ = c;
        check(pxg_set_layer_masks(w_, layers_.rows()), "layers");
        geometry_dirty_ = !hull_arrays_.empty() || !mesh_arrays_.empty();
        joints_dirty_ = !joints_.empty();
        ignored_dirty_ = !ignored_.empty();
        cloths_dirty_ = !cloths_.empty();
        strands_dirty_ = !strand_sets_.empty();
        pending_cache_keys_ = std::move(keys);
        pending_cache_imp_ = std::move(imp);
    }

public:

private:
    static void set3(float* d, Vec3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }
    uint32_t generations_at(uint32_t i) const { return i < generations_.size() ? generations_[i] : 0u; }
    pxg_joint joint_base(uint32_t type, BodyId a, BodyId b, Vec3 anchor_a, Vec3 anchor_b, bool collide_connected) {
        pull();
        pxg_joint j;
        std::memset(&j, 0, sizeof(j));
        j.type = type;
        j.body_a = a.index;
        j.body_b = b.index;
        j.valid = j.enabled = 1;
        j.collide_connected = collide_connected ? 1u : 0u;
        set3(j.local_anchor_a, anchor_a);
        set3(j.local_anchor_b, anchor_b);
        j.local_axis_a[2] = j.local_axis_b[2] = 1.0f;
        j.swing_limit = -1.0f;
        j.twist_min = 1.0f;
        j.twist_max = -1.0f;
        float ia[4];
        quat_ops::inverse(bodies_[a.index].orientation, ia);
        quat_ops::mul(ia, bodies_[b.index].orientation, j.rest_relative_rotation);
        return j;
    }
    JointId insert_joint(const pxg_joint& j) {  // world.h:1663-1679
        pull_joints();
        uint32_t index;
        if (!free_joints_.empty()) {
            index = free_joints_.back();
            free_joints_.pop_back();
            joints_[index] = j;
        } else {
            index = static_cast<uint32_t>(joints_.size());
            joints_.push_back(j);
            joint_generations_.push_back(0);
            joint_alive_.push_back(0);
        }
        joint_alive_[index] = 1;
        joints_dirty_ = true;
        return JointId{index, joint_generations_[index]};
    }
    void pull_joints() {
        if (!joints_on_device_ || joints_.empty() || joints_dirty_) return;
        check(pxg_get_joints(w_, joints_.data(), static_cast<uint32_t>(joints_.size())), "get_joint");
        joints_on_device_ = false;
    }
    void push_joints() {
        if (joints_dirty_) {
            check(pxg_set_joints(w_, joints_.data(), static_cast<uint32_t>(joints_.size())), "joints");
            joints_dirty_ = false;
        }
        if (ignored_dirty_) {
            std::vector<uint32_t> ab;
            for (const auto& p : ignored_) { ab.push_back(p.first); ab.push_back(p.second); }
            check(pxg_set_ignored_pairs(w_, ab.data(), static_cast<uint32_t>(ignored_.size())), "ignored pairs");
            ignored_dirty_ = false;
        }
    }

    bool alive_on_host(uint32_t i) const {
        // aliveness changes only on the host (add/remove), so the mirror's flag is current
        return bodies_[i].alive != 0;
    }

    static pxg_cast_query cast_base(float max_distance, Vec3 origin, Quat orientation, Vec3 dir) {
        pxg_cast_query c;
        std::memset(&c, 0, sizeof(c));
        c.shape = no_shape();
        c.shape.enabled = 1;
        set3(c.position, origin);
        c.orientation[0] = orientation.x;
        c.orientation[1] = orientation.y;
        c.orientation[2] = orientation.z;
        c.orientation[3] = orientation.w;
        set3(c.direction, dir);
        c.max_distance = max_distance;
        return c;
    }
    bool shape_cast(const pxg_cast_query& c, RaycastHit& hit, const QueryFilter& filter) {
        push();
        pxg_query_filter f = to_filter(filter);
        pxg_raycast_hit raw;
        check(pxg_shape_cast(w_, &c, &f, &raw), "shape_cast");
        if (!raw.hit) return false;
        hit = from_hit(raw);
        hit.started_inside = raw.started_inside != 0;
        return true;
    }
    static pxg_overlap_query overlap_base(uint32_t type) {
        pxg_overlap_query q;
        std::memset(&q, 0, sizeof(q));
        q.type = type;
        q.orientation[3] = 1.0f;
        return q;
    }
    std::vector<BodyId> overlap(const pxg_overlap_query& q, const QueryFilter& filter) {
        push();
        pxg_query_filter f = to_filter(filter);
        uint32_t n = 0;
        check(pxg_overlap(w_, &q, &f, nullptr, 0, &n), "overlap");
        std::vector<uint32_t> raw(n);
        if (n) check(pxg_overlap(w_, &q, &f, raw.data(), n, &n), "overlap");
        std::vector<BodyId> out;
        for (uint32_t i = 0; i < n && i < raw.size(); ++i)
            out.push_back(BodyId{raw[i], raw[i] < generations_.size() ? generations_[raw[i]] : 0u});
        return out;
    }
    static pxg_ray to_ray(const Ray& r) {
        pxg_ray o;
        std::memset(&o, 0, sizeof(o));
        set3(o.origin, r.origin);
        set3(o.direction, r.direction);
        o.max_distance = r.max_distance;
        return o;
    }
    pxg_query_filter to_filter(const QueryFilter& f) const {
        pxg_query_filter o;
        std::memset(&o, 0, sizeof(o));
        o.layer_mask = f.layer_mask;
        o.include_triggers = f.include_triggers ? 1u : 0u;
        o.ignore_body = is_valid(f.ignore) ? f.ignore.index : PXG_NO_BODY;
        return o;
    }
    RaycastHit from_hit(const pxg_raycast_hit& h) const {
        RaycastHit o;
        o.point = Vec3{h.point[0], h.point[1], h.point[2]};
        o.normal = Vec3{h.normal[0], h.normal[1], h.normal[2]};
        o.distance = h.distance;
        o.body = BodyId{h.body, h.body < generations_.size() ? generations_[h.body] : 0u};
        o.shape_index = h.shape_index;
        return o;
    }

    // device -> host mirror; the first caller downloads, concurrent callers wait for it
    void pull() {
        // between the two phases this is the collide phase's state (forces integrated)
        if (!state_on_device_.load(std::memory_order_acquire)) return;
        std::lock_guard<std::mutex> lock(pull_mu_);
        if (!state_on_device_.load(std::memory_order_relaxed)) return;
        if (!bodies_.empty()) {
            check(pxg_get_bodies(w_, bodies_.data(), static_cast<uint32_t>(bodies_.size())), "get_body");
            on_device_ = bodies_;
        }
        state_on_device_.store(false, std::memory_order_release);
    }
    // the write-back's poses: from the device after a step, else from the host mirror
    void pull_poses() {
        if (poses_valid_.load(std::memory_order_acquire)) return;
        std::lock_guard<std::mutex> lock(pull_mu_);
        if (poses_valid_.load(std::memory_order_relaxed)) return;
        poses_.resize(bodies_.size());
        if (state_on_device_.load(std::memory_order_relaxed)) {
            if (!poses_.empty()) check(pxg_get_poses(w_, poses_.data(), static_cast<uint32_t>(poses_.size())), "pose");
        } else {
            for (size_t i = 0; i < bodies_.size(); ++i) {
                const pxg_body& b = bodies_[i];
                pxg_pose& p = poses_[i];
                std::memcpy(p.position, b.position, sizeof(p.position));
                std::memcpy(p.orientation, b.orientation, sizeof(p.orientation));
                std::memcpy(p.prev_position, b.prev_position, sizeof(p.prev_position));
                std::memcpy(p.prev_orientation, b.prev_orientation, sizeof(p.prev_orientation));
                p.type = b.type;
                p.interpolation = b.interpolation;
            }
        }
        poses_valid_.store(true, std::memory_order_release);
    }
    // uploads the bodies whose host record differs from what the device holds (all of them when
    // most changed or a type changed)
    void upload_changed_bodies(const char* what) {
        const uint32_t n = static_cast<uint32_t>(bodies_.size());
        std::vector<uint32_t> idx;
        std::vector<pxg_body> rec;
        const bool known = on_device_.size() == bodies_.size();
        for (uint32_t i = 0; known && i < n; ++i) {
            if (std::memcmp(&bodies_[i], &on_device_[i], sizeof(pxg_body)) == 0) continue;
            idx.push_back(i);
            rec.push_back(bodies_[i]);
        }
        int r = PXG_ERR_STATE;
        if (known && idx.size() * 4 <= n)
            r = idx.empty() ? PXG_OK : pxg_set_bodies_subset(w_, idx.data(), rec.data(), static_cast<uint32_t>(idx.size()));
        if (r == PXG_OK) {
            bodies_uploaded_ += static_cast<uint32_t>(idx.size());
        } else {
            check(pxg_set_bodies(w_, bodies_.data(), n), what);
            bodies_uploaded_ += n;
        }
        on_device_ = bodies_;
        poses_valid_.store(false, std::memory_order_release);
    }

    // host -> device before a step
    void push() {
        bodies_uploaded_ = 0;
        if (std::memcmp(&applied_cfg_, &cfg_, sizeof(cfg_)) != 0) {
            check(pxg_set_config(w_, &cfg_), "config");
            applied_cfg_ = cfg_;
        }
        if (layers_.version() != applied_layers_) {
            check(pxg_set_layer_masks(w_, layers_.rows()), "layers");
            applied_layers_ = layers_.version();
        }
        const uint32_t n = static_cast<uint32_t>(bodies_.size());
        if (structure_dirty_) {
            // the broadphase keeps each surviving shape's fat AABB; new shapes get tight + margin
            std::vector<float> fat(size_t(uploaded_shapes_) * 6);
            if (uploaded_shapes_) check(pxg_get_fat_aabbs(w_, fat.data(), uploaded_shapes_), "fat AABBs");
            if (n > cfg_.max_bodies || n > cfg_.max_shapes) grow(n);
            check(pxg_set_bodies(w_, bodies_.data(), n), "bodies");
            bodies_uploaded_ = n;
            on_device_ = bodies_;
            poses_valid_.store(false, std::memory_order_release);
            if (geometry_dirty_) {
                // the old shapes may index past the new tables: clear them, then the tables, then the shapes
                check(pxg_set_shapes(w_, shapes_.data(), 0), "shapes");
                push_geometry();
            }
            check(pxg_set_shapes(w_, shapes_.data(), n), "shapes");
            std::vector<float> now(size_t(n) * 6);
            check(pxg_get_fat_aabbs(w_, now.data(), n), "fat AABBs");
            std::vector<char> fresh(n, 0);
            for (uint32_t i : new_shapes_) fresh[i] = 1;
            for (uint32_t i = 0; i < uploaded_shapes_ && i < n; ++i)
                if (!fresh[i]) std::memcpy(&now[size_t(i) * 6], &fat[size_t(i) * 6], 6 * sizeof(float));
            check(pxg_set_fat_aabbs(w_, now.data(), n), "fat AABBs");
            if (!pending_cache_keys_.empty()) {
                check(pxg_set_warm_cache(w_, pending_cache_keys_.data(), pending_cache_imp_.data(),
                                         static_cast<uint32_t>(pending_cache_keys_.size())), "warm cache");
                pending_cache_keys_.clear();
                pending_cache_imp_.clear();
            }
            uploaded_shapes_ = n;
            new_shapes_.clear();
            structure_dirty_ = false;
            bodies_dirty_.store(false, std::memory_order_relaxed);
        } else if (bodies_dirty_.load(std::memory_order_relaxed)) {
            upload_changed_bodies("bodies");
            bodies_dirty_.store(false, std::memory_order_relaxed);
        }
        push_joints();
        push_soft();
    }

    // ---- cloth and hair tables ----
    template <class T>
    static uint32_t take_slot(std::vector<T>& items, std::vector<uint32_t>& gens, std::vector<uint8_t>& alive,
                              std::vector<uint32_t>& free_list) {
        uint32_t index;
        if (!free_list.empty()) {
            index = free_list.back();
            free_list.pop_back();
        } else {
            index = static_cast<uint32_t>(items.size());
            items.emplace_back();
            gens.push_back(0);
            alive.push_back(0);
        }
        alive[index] = 1;
        return index;
    }
    static std::size_t live_count(const std::vector<uint8_t>& alive) {
        std::size_t n = 0;
        for (uint8_t a : alive) n += a ? 1 : 0;
        return n;
    }
    static pxg_cloth cloth_record(Cloth& c, bool alive) {
        pxg_cloth d;
        std::memset(&d, 0, sizeof(d));
        d.particles = c.particles.data();
        d.n_particles = static_cast<uint32_t>(c.particles.size());
        d.columns = c.columns;
        d.rows = c.rows;
        d.stretch = c.stretch.data();
        d.n_stretch = static_cast<uint32_t>(c.stretch.size());
        d.bend = c.bend.data();
        d.n_bend = static_cast<uint32_t>(c.bend.size());
        d.stretch_batches = c.stretch_batches.data();
        d.n_stretch_batches = static_cast<uint32_t>(c.stretch_batches.size() / 2);
        d.bend_batches = c.bend_batches.data();
        d.n_bend_batches = static_cast<uint32_t>(c.bend_batches.size() / 2);
        d.tethers = c.tethers.data();
        d.n_tethers = static_cast<uint32_t>(c.tethers.size());
        d.anchors = c.anchors.data();
        d.n_anchors = static_cast<uint32_t>(c.anchors.size());
        d.triangles = c.triangles.data();
        d.n_triangle_indices = static_cast<uint32_t>(c.triangles.size());
        d.params = c.params;
        d.alive = alive ? 1u : 0u;
        d.enabled = c.enabled ? 1u : 0u;
        d.awake = c.awake ? 1u : 0u;
        d.sleep_timer = c.sleep_timer;
        d.time = c.time;
        std::memcpy(d.bounds_min, c.bounds_min, sizeof(d.bounds_min));
        std::memcpy(d.bounds_max, c.bounds_max, sizeof(d.bounds_max));
        return d;
    }
    static pxg_strand_set strand_record(StrandSet& s, std::vector<uint32_t>& ids, bool alive) {
        pxg_strand_set d;
        std::memset(&d, 0, sizeof(d));
        d.particles = s.particles.data();
        d.n_particles = static_cast<uint32_t>(s.particles.size());
        d.stretch = s.stretch.data();
        d.n_stretch = static_cast<uint32_t>(s.stretch.size());
        d.stretch_batches = s.stretch_batches.data();
        d.n_stretch_batches = static_cast<uint32_t>(s.stretch_batches.size() / 2);
        d.rest_local = s.rest_local.data();
        d.rest_offset = s.rest_offset.data();
        d.strands = s.strands.data();
        d.n_strands = static_cast<uint32_t>(s.strands.size());
        ids.clear();
        for (const auto& b : s.collide_bodies) {
            ids.push_back(b.first);
            ids.push_back(b.second);
        }
        d.collide_bodies = ids.data();
        d.n_collide_bodies = static_cast<uint32_t>(s.collide_bodies.size());
        d.params = s.params;
        d.alive = alive ? 1u : 0u;
        d.enabled = s.enabled ? 1u : 0u;
        d.awake = s.awake ? 1u : 0u;
        d.sleep_timer = s.sleep_timer;
        d.time = s.time;
        std::memcpy(d.bounds_min, s.bounds_min, sizeof(d.bounds_min));
        std::memcpy(d.bounds_max, s.bounds_max, sizeof(d.bounds_max));
        return d;
    }
    void push_soft() {
        if (cloths_dirty_) {
            std::vector<pxg_cloth> recs;
            for (size_t i = 0; i < cloths_.size(); ++i) recs.push_back(cloth_record(cloths_[i], cloth_alive_[i] != 0));
            check(pxg_set_cloths(w_, recs.data(), static_cast<uint32_t>(recs.size())), "cloths");
            cloths_dirty_ = false;
        }
        if (strands_dirty_) {
            std::vector<std::vector<uint32_t>> ids(strand_sets_.size());
            std::vector<pxg_strand_set> recs;
            for (size_t i = 0; i < strand_sets_.size(); ++i)
                recs.push_back(strand_record(strand_sets_[i], ids[i], strand_alive_[i] != 0));
            check(pxg_set_strand_sets(w_, recs.data(), static_cast<uint32_t>(recs.size())), "strand sets");
            strands_dirty_ = false;
        }
    }
    // device -> host: the alive cloths' and grooms' state after a frame
    void pull_soft() {
        if (!soft_on_device_) return;
        std::vector<uint32_t> ids;
        for (uint32_t i = 0; i < cloths_.size(); ++i) {
            if (!cloth_alive_[i]) continue;
            Cloth& c = cloths_[i];
            pxg_cloth d = cloth_record(c, true);
            check(pxg_get_cloth(w_, i, &d), "get_cloth");
            c.awake = d.awake != 0;
            c.sleep_timer = d.sleep_timer;
            c.time = d.time;
            std::memcpy(c.bounds_min, d.bounds_min, sizeof(c.bounds_min));
            std::memcpy(c.bounds_max, d.bounds_max, sizeof(c.bounds_max));
        }
        for (uint32_t i = 0; i < strand_sets_.size(); ++i) {
            if (!strand_alive_[i]) continue;
            StrandSet& s = strand_sets_[i];
            pxg_strand_set d = strand_record(s, ids, true);
            check(pxg_get_strand_set(w_, i, &d), "get_strands");
            s.awake = d.awake != 0;
            s.sleep_timer = d.sleep_timer;
            s.time = d.time;
            std::memcpy(s.bounds_min, d.bounds_min, sizeof(s.bounds_min));
            std::memcpy(s.bounds_max, d.bounds_max, sizeof(s.bounds_max));
        }
        soft_on_device_ = false;
    }

    struct Deferred {
        bool create;
        BodyId id;
        pxg_body body;
        pxg_shape shape;
    };
    std::vector<Deferred> deferred_;
    pxg_world* w_ = nullptr;
    pxg_config cfg_, applied_cfg_;
    LayerMatrix layers_;
    uint64_t applied_layers_ = 0;
    std::vector<pxg_body> bodies_;
    std::vector<pxg_shape> shapes_;  // shape slot i belongs to body i
    std::vector<uint32_t> generations_, free_, new_shapes_;
    std::vector<uint64_t> pending_cache_keys_;
    std::vector<float> pending_cache_imp_;
    uint32_t uploaded_shapes_ = 0;
    bool structure_dirty_ = false, in_substep_ = false;
    std::atomic<bool> bodies_dirty_{false}, state_on_device_{false}, poses_valid_{false};
    std::mutex pull_mu_;
    std::vector<pxg_body> on_device_;  // the records the device holds, as of the last pull or upload
    std::vector<pxg_pose> poses_;
    uint32_t bodies_uploaded_ = 0;
    float interpolation_alpha_ = 0.0f;
    std::vector<pxg_joint> joints_;
    std::vector<uint32_t> joint_generations_, free_joints_;
    std::vector<uint8_t> joint_alive_;
    std::vector<std::pair<uint32_t, uint32_t>> ignored_;
    bool joints_dirty_ = false, ignored_dirty_ = false, joints_on_device_ = false;
    struct HullArrays {
        pxg_hull desc;
        std::vector<float> v, planes;
        std::vector<uint32_t> starts, fv, edges;
    };
    struct MeshArrays {
        std::vector<float> positions;
        std::vector<uint32_t> indices;
    };
    std::vector<HullArrays> hull_arrays_;
    std::vector<MeshArrays> mesh_arrays_;
    bool geometry_dirty_ = false;
    void push_geometry() {
        std::vector<pxg_hull> hs;
        for (HullArrays& a : hull_arrays_) {
            pxg_hull h = a.desc;
            h.vertices = a.v.data(); h.face_planes = a.planes.data(); h.face_starts = a.starts.data();
            h.face_vertices = a.fv.data(); h.edges = a.edges.empty() ? nullptr : a.edges.data();
            hs.push_back(h);
        }
        std::vector<pxg_mesh_desc> ms;
        for (MeshArrays& a : mesh_arrays_)
            ms.push_back(pxg_mesh_desc{a.positions.data(), static_cast<uint32_t>(a.positions.size() / 3),
                                       static_cast<uint32_t>(a.indices.size()), a.indices.data()});
        check(pxg_set_hulls(w_, hs.data(), static_cast<uint32_t>(hs.size())), "hulls");
        check(pxg_set_meshes(w_, ms.data(), static_cast<uint32_t>(ms.size())), "meshes");
        geometry_dirty_ = false;
    }
    float accumulator_ = 0.0f;
    std::vector<Cloth> cloths_;
    std::vector<StrandSet> strand_sets_;
    std::vector<uint32_t> cloth_generations_, free_cloths_, strand_generations_, free_strands_;
    std::vector<uint8_t> cloth_alive_, strand_alive_;
    bool cloths_dirty_ = false, strands_dirty_ = false, soft_on_device_ = false;
    mutable std::vector<pxg_manifold> manifolds_;
    mutable bool manifolds_valid_ = false;
    mutable std::vector<ContactEvent> events_;
    mutable bool events_valid_ = false;
};

}  // namespace pxg
