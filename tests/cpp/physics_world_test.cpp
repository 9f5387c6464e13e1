// The reference's PhysicsWorld tests (physxcoopa/tests/physics_world_test.cpp, sleep_test.cpp,
// contact_solver_test.cpp) written against the C++ facade (include/pxg_world.hpp), which steps on
// the GPU through the C ABI. Built in-tree by physxcoopa_amd/build.py; run by
// tests/test_gpu_cpp_facade.py. Prints one line per test and exits non-zero on any failure.
#include "pxg_world.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

namespace {

int g_failures = 0;
std::string g_current;

#define CHECK(cond)                                                                       \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);               \
            ++g_failures;                                                                 \
            return;                                                                       \
        }                                                                                 \
    } while (0)
#define CHECK_NEAR(a, b, tol)                                                             \
    do {                                                                                  \
        double a_ = (a), b_ = (b);                                                        \
        if (!(std::fabs(a_ - b_) <= (tol))) {                                             \
            std::printf("  FAILED %s:%d: %s = %.7g, expected %.7g +- %g\n", __FILE__, __LINE__, #a, a_, b_, \
                        double(tol));                                                     \
            ++g_failures;                                                                 \
            return;                                                                       \
        }                                                                                 \
    } while (0)

const float kH = 1.0f / 60.0f;

// dynamics::Body defaults (dynamics/body.h:85-125)
pxg_body body(uint32_t type, float x, float y, float z) {
    pxg_body b;
    std::memset(&b, 0, sizeof(b));
    b.type = type;
    b.awake = 1;
    b.use_gravity = 1;
    b.position[0] = x;
    b.position[1] = y;
    b.position[2] = z;
    b.orientation[3] = 1.0f;
    b.mass = 1.0f;
    b.inv_mass = type == PXG_DYNAMIC ? 1.0f : 0.0f;
    b.angular_drag = 0.05f;
    b.inv_inertia_local[0] = b.inv_inertia_local[1] = b.inv_inertia_local[2] = 1.0f;
    return b;
}
// collision::Shape::make_box / make_sphere (collision/shape.h:109-126) with the default material
pxg_shape box_shape(float hx, float hy, float hz) {
    pxg_shape s = pxg::no_shape();
    s.enabled = 1;
    s.type = PXG_BOX;
    s.half_extents[0] = hx;
    s.half_extents[1] = hy;
    s.half_extents[2] = hz;
    s.capsule_axis = 2;
    s.dynamic_friction = s.static_friction = 0.6f;
    return s;
}
pxg_shape sphere_shape(float r) {
    pxg_shape s = box_shape(0, 0, 0);
    s.type = PXG_SPHERE;
    s.radius = r;
    return s;
}

// tests/support/physics_fixtures.h:37-72
pxg::BodyId add_static_ground(pxg::PhysicsWorld& w, float restitution = 0.0f) {
    pxg_shape s = box_shape(4.0f, 4.0f, 0.1f);
    s.restitution = restitution;
    return w.add_body(body(PXG_STATIC, 0, 0, -0.1f), s);
}
pxg::BodyId add_dynamic_sphere(pxg::PhysicsWorld& w, float x, float y, float z, float r, float restitution = 0.0f) {
    pxg_body b = body(PXG_DYNAMIC, x, y, z);
    float inv = 1.0f / (0.4f * b.mass * r * r);  // solid sphere (dynamics/inertia.h:36-47)
    b.inv_inertia_local[0] = b.inv_inertia_local[1] = b.inv_inertia_local[2] = inv;
    pxg_shape s = sphere_shape(r);
    s.restitution = restitution;
    return w.add_body(b, s);
}
pxg::BodyId add_dynamic_box(pxg::PhysicsWorld& w, float x, float y, float z, float hx, float hy, float hz,
                            float mass = 1.0f) {
    pxg_body b = body(PXG_DYNAMIC, x, y, z);
    b.mass = mass;
    b.inv_mass = 1.0f / mass;
    float ix = mass / 3.0f * (hy * hy + hz * hz), iy = mass / 3.0f * (hx * hx + hz * hz),
          iz = mass / 3.0f * (hx * hx + hy * hy);
    b.inv_inertia_local[0] = 1.0f / ix;
    b.inv_inertia_local[1] = 1.0f / iy;
    b.inv_inertia_local[2] = 1.0f / iz;
    return w.add_body(b, box_shape(hx, hy, hz));
}
float vlen(const float* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
// max |axis.z| over the body's local axes (physics_fixtures.h:178-185)
float best_up_alignment(const float* q) {
    float x = q[0], y = q[1], z = q[2], w = q[3];
    float ax = 2.0f * (x * z - w * y), ay = 2.0f * (y * z + w * x), az = 1.0f - 2.0f * (x * x + y * y);
    return std::fmax(std::fabs(ax), std::fmax(std::fabs(ay), std::fabs(az)));
}

// physics_world_test.cpp:24-58
void free_fall_matches_analytic_while_static_and_kinematic_hold() {
    pxg::PhysicsWorld world;
    pxg::BodyId id = world.add_body(body(PXG_DYNAMIC, 0, 0, 0));
    pxg::BodyId kin = world.add_body(body(PXG_KINEMATIC, 1, 2, 3));
    pxg::BodyId stat = world.add_body(body(PXG_STATIC, 4, 5, 6));
    for (int i = 0; i < 60; ++i) world.step_fixed(kH);
    const pxg_body* b = world.get_body(id);
    CHECK(b != nullptr);
    float n = 60.0f;
    CHECK_NEAR(b->position[2], -9.81f * kH * kH * n * (n + 1.0f) * 0.5f, 0.01);
    CHECK_NEAR(b->position[2], -0.5f * 9.81f, 0.2);
    CHECK_NEAR(world.get_body(kin)->position[0], 1.0f, 1e-5);
    CHECK_NEAR(world.get_body(kin)->position[2], 3.0f, 1e-5);
    CHECK_NEAR(world.get_body(stat)->position[1], 5.0f, 1e-5);
}

// physics_world_test.cpp:60-69
void stale_body_handle_is_rejected_after_slot_reuse() {
    pxg::PhysicsWorld world;
    pxg::BodyId a = world.add_body(body(PXG_DYNAMIC, 0, 0, 0));
    world.remove_body(a);
    pxg::BodyId b = world.add_body(body(PXG_DYNAMIC, 0, 0, 0));
    CHECK(a.index == b.index);
    CHECK(a.generation != b.generation);
    CHECK(!world.is_valid(a));
    CHECK(world.is_valid(b));
    CHECK(world.get_body(a) == nullptr);
    world.remove_body(a);  // stale: a no-op
    CHECK(world.is_valid(b));
    world.step_fixed(kH);
    CHECK(world.get_body(b) != nullptr);
}

// sleep_test.cpp:17-34
void resting_sphere_falls_asleep_at_rest_height() {
    pxg::PhysicsWorld world;
    add_static_ground(world);
    pxg::BodyId id = add_dynamic_sphere(world, 0, 0, 2, 0.5f);
    bool slept = false;
    for (int i = 0; i < 400 && !slept; ++i) {
        world.step_fixed(kH);
        if (!world.get_body(id)->awake) slept = true;
    }
    CHECK(slept);
    const pxg_body* b = world.get_body(id);
    CHECK_NEAR(b->position[2], 0.5f, 0.05);
    CHECK(vlen(b->linear_velocity) < world.config().sleep_linear + 1e-4f);
}

// sleep_test.cpp:42-62
void moving_kinematic_platform_wakes_sleeping_body() {
    pxg::PhysicsWorld world;
    pxg::BodyId platform = world.add_body(body(PXG_KINEMATIC, 0, 0, -0.5f), box_shape(2, 2, 0.5f));
    pxg::BodyId sphere = add_dynamic_sphere(world, 0, 0, 1, 0.5f);
    for (int i = 0; i < 120; ++i) world.step_fixed(kH);
    CHECK(!world.get_body(sphere)->awake);
    world.get_body(platform)->linear_velocity[0] = 3.0f;
    world.step_fixed(kH);
    CHECK(world.get_body(sphere)->awake);
}

// sleep_test.cpp:71-95 (a body added mid-run: the structural edit path)
void spinning_kinematic_body_wakes_sleeping_body_it_sweeps() {
    pxg::PhysicsWorld world;
    add_static_ground(world);
    pxg::BodyId sphere = add_dynamic_sphere(world, 3, 0, 1, 0.5f);
    for (int i = 0; i < 120; ++i) world.step_fixed(kH);
    CHECK(!world.get_body(sphere)->awake);
    const pxg_body* s = world.get_body(sphere);
    pxg_body paddle = body(PXG_KINEMATIC, s->position[0], s->position[1], s->position[2]);
    paddle.angular_velocity[2] = 2.0f;
    world.add_body(paddle, box_shape(1, 1, 1));
    world.step_fixed(kH);
    CHECK(world.get_body(sphere)->awake);
}

// sleep_test.cpp:103-124
void sleeping_body_ignores_awake_neighbours_impulse() {
    pxg::PhysicsWorld world;
    world.set_gravity(0, 0, 0);
    pxg::BodyId sleeper = add_dynamic_sphere(world, 0, 0, 0, 0.5f);
    pxg_body* sb = world.get_body(sleeper);
    sb->awake = 0;
    pxg::BodyId mover = add_dynamic_sphere(world, 0.9f, 0, 0, 0.5f);
    world.get_body(mover)->linear_velocity[0] = -2.0f;
    world.step_fixed(kH);
    const pxg_body* s = world.get_body(sleeper);
    CHECK(!s->awake);
    CHECK(vlen(s->linear_velocity) < 1e-6f);
    CHECK(vlen(s->angular_velocity) < 1e-6f);
}

// contact_solver_test.cpp:19-61
void restitution_one_bounces_high_and_zero_never_bounces() {
    {
        pxg::PhysicsWorld world;
        add_static_ground(world, 1.0f);
        pxg::BodyId id = add_dynamic_sphere(world, 0, 0, 2, 0.5f, 1.0f);
        bool bounced = false;
        float max_z = 0.0f;
        for (int i = 0; i < 300; ++i) {
            world.step_fixed(kH);
            const pxg_body* b = world.get_body(id);
            if (b->linear_velocity[2] > 0.0f) {
                bounced = true;
                max_z = std::fmax(max_z, b->position[2]);
            }
        }
        CHECK(bounced);
        CHECK(max_z > 0.5f + 0.6f * 1.5f);
    }
    {
        pxg::PhysicsWorld world;
        add_static_ground(world);
        pxg::BodyId id = add_dynamic_sphere(world, 0, 0, 2, 0.5f);
        bool contacted = false;
        float max_up = 0.0f;
        for (int i = 0; i < 300; ++i) {
            world.step_fixed(kH);
            const pxg_body* b = world.get_body(id);
            if (b->position[2] < 0.6f) contacted = true;
            if (contacted) max_up = std::fmax(max_up, b->linear_velocity[2]);
        }
        CHECK(contacted);
        CHECK(max_up < 1.0f);
    }
}

// contact_solver_test.cpp:120-150
void fast_box_does_not_tunnel_through_thin_wall() {
    pxg::PhysicsWorld world;
    world.set_gravity(0, 0, 0);
    world.add_body(body(PXG_STATIC, 5, 0, 0), box_shape(0.05f, 2, 2));
    pxg::BodyId box = add_dynamic_box(world, 4.5f, 0, 0, 0.1f, 0.1f, 0.1f);
    world.get_body(box)->linear_velocity[0] = 50.0f;
    world.step_fixed(kH);
    const pxg_body* b = world.get_body(box);
    CHECK(b->position[0] < 5.0f);
    CHECK(b->linear_velocity[0] < 50.0f);
    for (int i = 0; i < 300; ++i) world.step_fixed(kH);
    b = world.get_body(box);
    CHECK(b->position[0] < 4.951f);
    CHECK(vlen(b->linear_velocity) < 0.05f);
}

// contact_solver_test.cpp:152-172
void ten_box_stack_settles_flat_and_sleeps() {
    pxg::PhysicsWorld world;
    world.add_body(body(PXG_STATIC, 0, 0, -0.5f), box_shape(5, 5, 0.5f));
    std::vector<pxg::BodyId> ids;
    for (int i = 0; i < 10; ++i) ids.push_back(add_dynamic_box(world, 0, 0, 0.5f * (2 * i + 1), 0.5f, 0.5f, 0.5f));
    for (int i = 0; i < 1000; ++i) world.step_fixed(kH);
    for (auto id : ids) {
        const pxg_body* b = world.get_body(id);
        CHECK(std::fabs(b->position[0]) < 0.01f);
        CHECK(std::fabs(b->position[1]) < 0.01f);
        CHECK(best_up_alignment(b->orientation) > 0.999f);
        CHECK(!b->awake);
    }
}

// contact_solver_test.cpp:204-237
void mass_ratio_100_to_1_stack_stays_stable() {
    pxg::PhysicsWorld world;
    world.add_body(body(PXG_STATIC, 0, 0, -0.5f), box_shape(5, 5, 0.5f));
    pxg::BodyId light = add_dynamic_box(world, 0, 0, 0.5f, 0.5f, 0.5f, 0.5f, 1.0f);
    pxg::BodyId heavy = add_dynamic_box(world, 0, 0, 1.5f, 0.5f, 0.5f, 0.5f, 100.0f);
    for (int i = 0; i < 500; ++i) {
        world.step_fixed(kH);
        const pxg_body* l = world.get_body(light);
        const pxg_body* h = world.get_body(heavy);
        CHECK(std::isfinite(l->position[0]) && std::isfinite(l->position[2]) && std::isfinite(h->position[2]));
        CHECK(vlen(l->linear_velocity) < 50.0f && vlen(h->linear_velocity) < 50.0f);
    }
    CHECK(world.get_body(light)->position[2] > 0.3f && world.get_body(light)->position[2] < 0.7f);
    CHECK(world.get_body(heavy)->position[2] > 1.3f && world.get_body(heavy)->position[2] < 1.7f);
}

// physics_world_test.cpp:177-207: the same scene stepped twice hashes the same
void same_scene_twice_hashes_the_same() {
    uint64_t hashes[2];
    for (int run = 0; run < 2; ++run) {
        pxg::PhysicsWorld world;
        add_static_ground(world);
        uint32_t seed = 1234u;
        auto rnd = [&](float lo, float hi) {
            seed = seed * 1664525u + 1013904223u;
            return lo + (hi - lo) * float(seed >> 8) / float(1u << 24);
        };
        for (int i = 0; i < 80; ++i) add_dynamic_sphere(world, rnd(-3, 3), rnd(-3, 3), rnd(2, 8), 0.4f);
        for (int i = 0; i < 200; ++i) world.step_fixed(kH);
        hashes[run] = world.world_state_hash();
    }
    CHECK(hashes[0] == hashes[1]);
}

// world.h:70-78: a slot sees the collide phase's state and its velocity change reaches the solve;
// add_body from a slot is deferred to the next step and returns an invalid handle
void on_substep_slot_changes_velocity_and_defers_creation() {
    pxg::PhysicsWorld world;
    world.set_gravity(0, 0, 0);
    pxg::BodyId id = world.add_body(body(PXG_DYNAMIC, 0, 0, 0));
    pxg::BodyId created;
    int calls = 0;
    world.on_substep.connect([&](pxg::PhysicsWorld& w, float h) {
        CHECK(h == kH);
        if (calls++ == 0) {
            w.get_body(id)->linear_velocity[0] = 6.0f;
            created = w.add_body(body(PXG_DYNAMIC, 10, 0, 0));
        }
    });
    world.step_fixed(kH);
    CHECK(!created.is_valid());
    CHECK_NEAR(world.get_body(id)->position[0], 6.0f * kH, 1e-6);
    CHECK(world.body_count() == 1);
    world.step_fixed(kH);
    CHECK(world.body_count() == 2);
    CHECK(calls == 2);
}


// --- joints (joints_test.cpp, sleep_test.cpp:134-175, world.h:378-536) ---
struct V3 {
    float x, y, z;
};
V3 v3(const float* p) { return V3{p[0], p[1], p[2]}; }
V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float length(V3 a) { return std::sqrt(dot(a, a)); }
V3 rotate(const float* q, const float* v) {
    float r[3];
    pxg::quat_ops::rotate(q, v, r);
    return V3{r[0], r[1], r[2]};
}
V3 world_anchor(const pxg_body& b, const float* local) {
    V3 r = rotate(b.orientation, local);
    return V3{b.position[0] + r.x, b.position[1] + r.y, b.position[2] + r.z};
}
// Twist of b relative to a about a's local axis, measured from the rest pose
// (hinge_current_angle, solver.h:357-368).
float twist_angle(const pxg_body& a, const pxg_body& b, const pxg_joint& j) {
    float ia[4], rel[4], ir[4], d[4];
    pxg::quat_ops::inverse(a.orientation, ia);
    pxg::quat_ops::mul(ia, b.orientation, rel);
    pxg::quat_ops::inverse(j.rest_relative_rotation, ir);
    pxg::quat_ops::mul(rel, ir, d);
    V3 ax = v3(j.local_axis_a);
    const float n = length(ax);
    ax = V3{ax.x / n, ax.y / n, ax.z / n};
    V3 dv{d[0], d[1], d[2]};
    float w = d[3] < -1.0f ? -1.0f : (d[3] > 1.0f ? 1.0f : d[3]);
    if (w < 0.0f) {
        w = -w;
        dv = V3{-dv.x, -dv.y, -dv.z};
    }
    return 2.0f * std::atan2(dot(dv, ax), w);
}
pxg_body capsule_body(float x, float y, float z, float r, float hh, float mass) {
    pxg_body b = body(PXG_DYNAMIC, x, y, z);
    b.mass = mass;
    b.inv_mass = 1.0f / mass;
    // inertia_for_shape for a Z capsule (dynamics/inertia.h:87-107)
    const float pi = 3.14159265358979f;
    const float vc = pi * r * r * (2.0f * hh), vh = (4.0f / 3.0f) * pi * r * r * r;
    const float mc = mass * vc / (vc + vh), mh = mass * vh / (vc + vh);
    const float axial = 0.5f * mc * r * r + 0.4f * mh * r * r;
    const float perp = mc * (3.0f * r * r + 4.0f * hh * hh) / 12.0f + 0.4f * mh * r * r + mh * hh * hh;
    b.inv_inertia_local[0] = b.inv_inertia_local[1] = 1.0f / perp;
    b.inv_inertia_local[2] = 1.0f / axial;
    return b;
}
pxg_shape capsule_shape(float r, float hh) {
    pxg_shape s = box_shape(0, 0, 0);
    s.type = PXG_CAPSULE;
    s.capsule_radius = r;
    s.capsule_half_height = hh;
    return s;
}
const float kPi = 3.14159265358979f;
float radians(float deg) { return deg * kPi / 180.0f; }

void hinge_door_swings_and_stops_at_its_limit() {
    pxg::PhysicsWorld world;
    world.set_gravity(0, 0, 0);
    pxg::BodyId frame = world.add_body(body(PXG_STATIC, 0, 0, 0));
    pxg_body d = body(PXG_DYNAMIC, 1, 0, 0);
    d.angular_velocity[2] = 5.0f;
    pxg::BodyId door = world.add_body(d);
    pxg::JointId joint = world.add_hinge_joint(frame, door, {0, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, 1}, true, 0.0f,
                                               radians(90.0f));
    CHECK(joint.is_valid());
    for (int i = 0; i < 300; ++i) world.step_fixed(kH);
    const pxg_joint* j = world.get_joint(joint);
    CHECK(j != nullptr);
    const pxg_body& a = *world.get_body(frame);
    const pxg_body& b = *world.get_body(door);
    CHECK_NEAR(length(sub(world_anchor(b, j->local_anchor_b), world_anchor(a, j->local_anchor_a))), 0.0, 0.02);
    CHECK_NEAR(rotate(b.orientation, j->local_axis_b).z, 1.0, 0.02);
    CHECK_NEAR(twist_angle(a, b, *j), radians(90.0f), radians(3.0f));
    CHECK(length(v3(b.angular_velocity)) < 0.05f);
    CHECK(j->limit_impulse != 0.0f || j->point_impulse[0] != 0.0f || j->point_impulse[1] != 0.0f);
}

void hinge_free_pendulum_swings_without_drift() {
    pxg::PhysicsWorld world;
    pxg::BodyId ceiling = world.add_body(body(PXG_STATIC, 0, 0, 0));
    pxg::BodyId pend = world.add_body(body(PXG_DYNAMIC, 1, 0, 0));
    pxg::JointId joint = world.add_hinge_joint(ceiling, pend, {0, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, 1, 0});
    CHECK(joint.is_valid());
    float max_drift = 0.0f, min_align = 1.0f;
    const float start_z = world.get_body(pend)->position[2];
    for (int i = 0; i < 120; ++i) {
        world.step_fixed(kH);
        const pxg_joint& j = *world.get_joint(joint);
        const pxg_body& a = *world.get_body(ceiling);
        const pxg_body& b = *world.get_body(pend);
        max_drift = std::max(max_drift, length(sub(world_anchor(b, j.local_anchor_b), world_anchor(a, j.local_anchor_a))));
        min_align = std::min(min_align, dot(rotate(a.orientation, j.local_axis_a), rotate(b.orientation, j.local_axis_b)));
    }
    CHECK(max_drift < 0.05f);
    CHECK(min_align > 0.98f);
    CHECK(world.get_body(pend)->position[2] < start_z - 0.3f);
}

void hinge_zero_range_limit_acts_rigid() {
    pxg::PhysicsWorld world;
    pxg::BodyId frame = world.add_body(body(PXG_STATIC, 0, 0, 0));
    pxg_body p = body(PXG_DYNAMIC, 1, 0, 0);
    p.angular_velocity[2] = 5.0f;
    pxg::BodyId panel = world.add_body(p);
    pxg::JointId joint = world.add_hinge_joint(frame, panel, {0, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, 1}, true, 0.0f, 0.0f);
    CHECK(joint.is_valid());
    for (int i = 0; i < 120; ++i) world.step_fixed(kH);
    const pxg_joint& j = *world.get_joint(joint);
    const pxg_body& a = *world.get_body(frame);
    const pxg_body& b = *world.get_body(panel);
    CHECK_NEAR(twist_angle(a, b, j), 0.0, radians(3.0f));
    CHECK_NEAR(length(sub(world_anchor(b, j.local_anchor_b), world_anchor(a, j.local_anchor_a))), 0.0, 0.05);
}

void ball_joint_holds_anchor_and_releases_on_remove() {
    pxg::PhysicsWorld world;
    pxg::BodyId ceiling = world.add_body(body(PXG_STATIC, 0, 0, 0));
    pxg_body bb = capsule_body(0, 0, -0.5f, 0.08f, 0.42f, 2.0f);
    bb.linear_velocity[0] = 2.0f;
    bb.linear_velocity[1] = 1.5f;
    pxg::BodyId bob = world.add_body(bb, capsule_shape(0.08f, 0.42f));
    pxg::JointId joint = world.add_ball_joint(ceiling, bob, {0, 0, 0}, {0, 0, 0.5f});
    CHECK(joint.is_valid());
    CHECK(world.get_joint(joint)->type == PXG_BALL);
    float max_drift = 0.0f, max_x = 0.0f, max_y = 0.0f;
    for (int i = 0; i < 180; ++i) {
        world.step_fixed(kH);
        const pxg_body& b = *world.get_body(bob);
        const pxg_joint& j = *world.get_joint(joint);
        max_drift = std::max(max_drift, length(world_anchor(b, j.local_anchor_b)));
        max_x = std::max(max_x, std::fabs(b.position[0]));
        max_y = std::max(max_y, std::fabs(b.position[1]));
    }
    CHECK(max_drift < 0.02f);
    CHECK(max_x > 0.15f && max_y > 0.1f);
    CHECK_NEAR(length(v3(world.get_body(bob)->position)), 0.5, 0.03);
    world.remove_joint(joint);
    CHECK(!world.is_valid(joint));
    CHECK(world.get_joint(joint) == nullptr);
    for (int i = 0; i < 30; ++i) world.step_fixed(kH);
    CHECK(length(v3(world.get_body(bob)->position)) > 0.8f);
}

void cone_twist_respects_swing_and_twist_limits() {
    pxg::PhysicsWorld world;
    world.set_gravity(0, 0, 0);
    pxg::BodyId shoulder = world.add_body(body(PXG_STATIC, 0, 0, 0));
    pxg_body ab = capsule_body(0, 0, -0.3f, 0.06f, 0.24f, 2.0f);
    ab.angular_velocity[0] = 6.0f;
    ab.angular_velocity[2] = 9.0f;
    pxg::BodyId arm = world.add_body(ab, capsule_shape(0.06f, 0.24f));
    const float swing_limit = radians(40.0f), twist_min = radians(-20.0f), twist_max = radians(25.0f);
    pxg::JointId joint = world.add_cone_twist_joint(shoulder, arm, {0, 0, 0}, {0, 0, 0.3f}, {0, 0, -1}, swing_limit,
                                                    twist_min, twist_max);
    CHECK(joint.is_valid());
    float max_swing = 0.0f, max_twist = -10.0f, min_twist = 10.0f, max_drift = 0.0f;
    for (int i = 0; i < 240; ++i) {
        world.step_fixed(kH);
        const pxg_body& a = *world.get_body(shoulder);
        const pxg_body& b = *world.get_body(arm);
        const pxg_joint& j = *world.get_joint(joint);
        V3 ta = rotate(a.orientation, j.local_axis_a), tb = rotate(b.orientation, j.local_axis_b);
        max_swing = std::max(max_swing, std::atan2(length(cross(ta, tb)), dot(ta, tb)));
        const float tw = twist_angle(a, b, j);
        max_twist = std::max(max_twist, tw);
        min_twist = std::min(min_twist, tw);
        max_drift = std::max(max_drift, length(world_anchor(b, j.local_anchor_b)));
    }
    const float tol = radians(4.0f);
    CHECK(max_swing <= swing_limit + tol);
    CHECK(max_swing >= swing_limit - tol);
    CHECK(min_twist >= twist_min - tol);
    CHECK(min_twist <= twist_min + tol);
    CHECK(max_twist <= twist_max + tol);
    CHECK(max_drift < 0.02f);
}

void joint_connected_bodies_share_a_sleep_island() {
    pxg::PhysicsWorld world;
    world.set_gravity(0, 0, 0);
    pxg::BodyId a = world.add_body(body(PXG_DYNAMIC, 0, 0, 0));
    pxg_body bb = body(PXG_DYNAMIC, 1, 0, 0);
    bb.angular_velocity[2] = 2.0f;
    pxg::BodyId b = world.add_body(bb);
    pxg::JointId joint = world.add_hinge_joint(a, b, {0, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, 1});
    CHECK(joint.is_valid());
    for (int i = 0; i < 90; ++i) world.step_fixed(kH);
    CHECK(length(v3(world.get_body(b)->angular_velocity)) > 0.5f);
    CHECK(world.get_body(a)->awake);
}

// remove_body turns every joint on the body off at once (world.h:130-139); the joint slot itself
// lives on until remove_joint, after which its stale handle is rejected when the slot is reused
void removing_a_body_invalidates_its_joints_and_stale_joint_handles_are_rejected() {
    pxg::PhysicsWorld world;
    pxg::BodyId ceiling = world.add_body(body(PXG_STATIC, 0, 0, 0));
    pxg::BodyId bob = world.add_body(body(PXG_DYNAMIC, 0, 0, -0.5f));
    pxg_body ob = body(PXG_DYNAMIC, 3, 0, -0.5f);
    ob.linear_velocity[1] = 1.0f;  // keeps swinging, so it never sleeps while held
    pxg::BodyId other = world.add_body(ob);
    auto radius = [&]() {
        const float* p = world.get_body(other)->position;
        return length(V3{p[0] - 3.0f, p[1], p[2]});
    };
    pxg::JointId j1 = world.add_ball_joint(ceiling, bob, {0, 0, 0}, {0, 0, 0.5f});
    pxg::JointId j2 = world.add_ball_joint(ceiling, other, {3, 0, 0}, {0, 0, 0.5f});
    for (int i = 0; i < 10; ++i) world.step_fixed(kH);
    CHECK_NEAR(radius(), 0.5, 0.03);
    world.remove_body(bob);
    CHECK(world.is_valid(j1));
    CHECK(world.get_joint(j1)->valid == 0);
    CHECK(world.get_joint(j2)->valid == 1);
    for (int i = 0; i < 30; ++i) world.step_fixed(kH);
    CHECK_NEAR(radius(), 0.5, 0.03);  // j2 still holds
    world.remove_joint(j1);
    pxg::JointId j3 = world.add_ball_joint(ceiling, other, {3, 0, 0}, {0, 0, 0.5f});
    CHECK(j3.index == j1.index);
    CHECK(!world.is_valid(j1));
    CHECK(world.is_valid(j3));
    CHECK(world.get_joint(j1) == nullptr);
    world.remove_joint(j2);
    world.remove_joint(j3);
    for (int i = 0; i < 30; ++i) world.step_fixed(kH);
    CHECK(world.get_body(other)->position[2] < -0.7f);  // no joint left: it falls
}

// set_pair_collision(a, b, false) drops the pair's contacts regardless of layers; turning it back
// on restores them; pair_collides reports the setting (world.h:527-536)
void pair_collision_switch_lets_a_sphere_fall_through_the_ground() {
    pxg::PhysicsWorld world;
    pxg::BodyId ground = add_static_ground(world);
    pxg::BodyId kept = world.add_body(body(PXG_DYNAMIC, 1, 0, 0.5f), sphere_shape(0.5f));
    pxg::BodyId dropped = world.add_body(body(PXG_DYNAMIC, -1, 0, 0.5f), sphere_shape(0.5f));
    CHECK(world.pair_collides(ground, dropped));
    world.set_pair_collision(ground, dropped, false);
    world.set_pair_collision(dropped, ground, false);  // same pair either way round
    CHECK(!world.pair_collides(dropped, ground));
    for (int i = 0; i < 60; ++i) world.step_fixed(kH);
    CHECK_NEAR(world.get_body(kept)->position[2], 0.5, 0.02);
    CHECK(world.get_body(dropped)->position[2] < -1.0f);
    world.set_pair_collision(ground, dropped, true);
    CHECK(world.pair_collides(ground, dropped));
    pxg::BodyId kept2 = world.add_body(body(PXG_DYNAMIC, 2, 2, 0.5f), sphere_shape(0.5f));
    world.set_pair_collision(ground, kept2, false);
    world.remove_body(kept2);  // a removed body forgets its pairs
    CHECK(world.pair_collides(ground, kept2));
}

}  // namespace

// contact_events_test.cpp:23-61: a 0.1 m probe flies through a static trigger sphere of radius 1
// with gravity off; one enter, one exit, a stay every step in between
void trigger_fires_enter_and_exit_once_with_stays_between() {
    pxg::PhysicsWorld world;
    world.set_gravity(0.0f, 0.0f, 0.0f);
    pxg_shape trig = sphere_shape(1.0f);
    trig.is_trigger = 1;
    pxg::BodyId trigger_id = world.add_body(body(PXG_STATIC, 0, 0, 0), trig);
    pxg_body probe = body(PXG_DYNAMIC, -3.0f, 0, 0);
    probe.linear_velocity[0] = 2.0f;
    float inv = 1.0f / (0.4f * 1.0f * 0.1f * 0.1f);
    probe.inv_inertia_local[0] = probe.inv_inertia_local[1] = probe.inv_inertia_local[2] = inv;
    pxg::BodyId probe_id = world.add_body(probe, sphere_shape(0.1f));
    int enter = 0, stay = 0, exit_ = 0;
    for (int i = 0; i < 200; ++i) {
        world.clear_events();
        world.step_fixed(kH);
        for (const pxg::ContactEvent& ev : world.events()) {
            if (!ev.is_trigger) continue;
            bool pair = (ev.a == trigger_id && ev.b == probe_id) || (ev.a == probe_id && ev.b == trigger_id);
            if (!pair) continue;
            if (ev.type == pxg::ContactEvent::Type::Enter) ++enter;
            else if (ev.type == pxg::ContactEvent::Type::Stay) ++stay;
            else ++exit_;
        }
    }
    CHECK(enter == 1);
    CHECK(exit_ == 1);
    CHECK(stay > 0);
}

// the sphere bounces off the ground (restitution 0.9 and 0, Average: 0.45): solid enters and
// exits; step() drains the queue each call (world.h:1230)
void bouncing_sphere_enters_and_exits_and_step_clears_the_queue() {
    pxg::PhysicsWorld world;
    pxg::BodyId ground = add_static_ground(world);
    pxg::BodyId ball = add_dynamic_sphere(world, 0, 0, 3.0f, 0.5f, 0.9f);
    int enter = 0, exit_ = 0;
    size_t most = 0;
    for (int i = 0; i < 300; ++i) {
        world.step(kH);
        most = std::max(most, world.events().size());
        for (const pxg::ContactEvent& ev : world.events()) {
            bool pair = (ev.a == ground && ev.b == ball) || (ev.a == ball && ev.b == ground);
            if (!pair || ev.is_trigger) continue;
            if (ev.type == pxg::ContactEvent::Type::Enter) ++enter;
            if (ev.type == pxg::ContactEvent::Type::Exit) ++exit_;
        }
    }
    CHECK(enter >= 1);
    CHECK(exit_ >= 1);
    CHECK(most <= 2);  // one substep per step(kH): at most this step's events of the one pair
}

// queries_test.cpp:22-37, 73-97, 270-293: exact distance, point and normal; every hit nearest
// first; any hit; the ignore filter
void raycasts_hit_boxes_with_exact_distances_and_filters() {
    pxg::PhysicsWorld world;
    pxg::BodyId b9 = world.add_body(body(PXG_STATIC, 9.0f, 0, 0), box_shape(0.5f, 0.5f, 0.5f));
    pxg::BodyId b3 = world.add_body(body(PXG_STATIC, 3.0f, 0, 0), box_shape(0.5f, 0.5f, 0.5f));
    pxg::BodyId b6 = world.add_body(body(PXG_STATIC, 6.0f, 0, 0), box_shape(0.5f, 0.5f, 0.5f));
    (void)b9;
    pxg::Ray ray;
    ray.origin = pxg::Vec3{0.0f, 0.0f, 0.0f};
    ray.direction = pxg::Vec3{1.0f, 0.0f, 0.0f};
    ray.max_distance = 100.0f;
    pxg::RaycastHit hit;
    CHECK(world.raycast(ray, hit));
    CHECK_NEAR(hit.distance, 2.5f, 1e-3f);
    CHECK_NEAR(hit.point.x, 2.5f, 1e-3f);
    CHECK_NEAR(hit.normal.x, -1.0f, 1e-3f);
    CHECK(hit.body == b3);
    std::vector<pxg::RaycastHit> all = world.raycast_all(ray);
    CHECK(all.size() == 3);
    CHECK_NEAR(all[0].distance, 2.5f, 1e-3f);
    CHECK_NEAR(all[1].distance, 5.5f, 1e-3f);
    CHECK_NEAR(all[2].distance, 8.5f, 1e-3f);
    CHECK(world.raycast_any(ray));
    pxg::Ray back = ray;
    back.direction = pxg::Vec3{-1.0f, 0.0f, 0.0f};
    CHECK(!world.raycast_any(back));
    pxg::QueryFilter ignore_near;
    ignore_near.ignore = b3;
    CHECK(world.raycast(ray, hit, ignore_near));
    CHECK(hit.body == b6);
    CHECK_NEAR(hit.distance, 5.5f, 1e-4f);
    CHECK(world.raycast_all(ray, ignore_near).size() == 2);
}

// queries_test.cpp:100-123: overlap queries find only the touched body
void overlaps_find_only_touched_bodies() {
    pxg::PhysicsWorld world;
    pxg::BodyId near_id = world.add_body(body(PXG_STATIC, 0, 0, 0), box_shape(0.5f, 0.5f, 0.5f));
    world.add_body(body(PXG_STATIC, 20.0f, 0, 0), box_shape(0.5f, 0.5f, 0.5f));
    std::vector<pxg::BodyId> s = world.overlap_sphere(pxg::Vec3{0.2f, 0.0f, 0.0f}, 0.5f);
    CHECK(s.size() == 1 && s[0] == near_id);
    std::vector<pxg::BodyId> b = world.overlap_box(pxg::Vec3{0, 0, 0}, pxg::Vec3{1, 1, 1}, pxg::Quat{});
    CHECK(b.size() == 1 && b[0] == near_id);
    std::vector<pxg::BodyId> c = world.overlap_capsule(pxg::Vec3{0, 0, -0.4f}, pxg::Vec3{0, 0, 0.4f}, 0.3f);
    CHECK(c.size() == 1 && c[0] == near_id);
}

// queries_test.cpp:132-174: a sphere cast meets a box corner at 4.5 - sqrt(0.07); a capsule
// cast meets a Y-axis capsule at 2.6
void sweeps_are_exact_against_box_corner_and_capsule() {
    pxg::PhysicsWorld world;
    pxg::BodyId box_id = world.add_body(body(PXG_STATIC, 5.0f, 0, 0), box_shape(0.5f, 0.5f, 0.5f));
    pxg::RaycastHit hit;
    CHECK(world.sphere_cast(pxg::Vec3{0.0f, 0.8f, 0.8f}, 0.5f, pxg::Vec3{1, 0, 0}, 100.0f, hit));
    CHECK_NEAR(hit.distance, 4.5f - std::sqrt(0.07f), 1e-3f);
    CHECK_NEAR(hit.point.x, 4.5f, 1e-3f);
    CHECK(hit.body == box_id && !hit.started_inside);
    // a box yawed 45 degrees meets the face x = 4.5 with its leading edge at 4.5 - 0.5 sqrt(2)
    pxg::Quat yaw{0.0f, 0.0f, std::sin(0.3926991f), std::cos(0.3926991f)};
    CHECK(world.box_cast(pxg::Vec3{0, 0, 0}, pxg::Vec3{0.5f, 0.5f, 0.5f}, yaw, pxg::Vec3{1, 0, 0}, 100.0f, hit));
    CHECK_NEAR(hit.distance, 4.5f - 0.5f * std::sqrt(2.0f), 1e-3f);
    CHECK_NEAR(hit.normal.x, -1.0f, 1e-3f);
    pxg::PhysicsWorld capsules;
    pxg_shape cap = pxg::no_shape();
    cap.enabled = 1;
    cap.type = PXG_CAPSULE;
    cap.capsule_radius = 0.2f;
    cap.capsule_half_height = 1.0f;
    cap.capsule_axis = 1;
    capsules.add_body(body(PXG_STATIC, 3.0f, 0.0f, 0.8f), cap);
    CHECK(capsules.capsule_cast(pxg::Vec3{0, 0, 0}, 0.3f, 0.5f, 2, pxg::Quat{}, pxg::Vec3{1, 0, 0}, 100.0f, hit));
    CHECK_NEAR(hit.distance, 2.6f, 1e-3f);
    CHECK_NEAR(hit.normal.x, -0.8f, 1e-2f);
    CHECK(!capsules.capsule_cast(pxg::Vec3{0, 0, -0.65f}, 0.3f, 0.5f, 2, pxg::Quat{}, pxg::Vec3{1, 0, 0}, 100.0f, hit));
}

// queries_test.cpp:335-360 (the convex part): push-outs from a box floor, exact depths and +Z normals
void compute_penetration_is_exact_against_box() {
    pxg::PhysicsWorld world;
    pxg::BodyId floor_id = world.add_body(body(PXG_STATIC, 0.0f, 0.0f, -0.5f), box_shape(2.0f, 2.0f, 0.5f));
    pxg_shape sphere = pxg::no_shape();
    sphere.enabled = 1;
    sphere.type = PXG_SPHERE;
    sphere.radius = 0.5f;
    std::vector<pxg::Penetration> p = world.compute_penetration(sphere, pxg::Vec3{0.0f, 0.0f, 0.4f}, pxg::Quat{});
    CHECK(p.size() == 1);
    if (p.size() == 1) {
        CHECK_NEAR(p[0].depth, 0.1f, 1e-4f);
        CHECK_NEAR(p[0].normal.z, 1.0f, 1e-4f);
        CHECK(p[0].body == floor_id && p[0].shape_index == 0u);
    }
    p = world.compute_penetration(box_shape(0.25f, 0.25f, 0.25f), pxg::Vec3{0.5f, 0.0f, 0.1f}, pxg::Quat{});
    CHECK(p.size() == 1);
    if (p.size() == 1) {
        CHECK_NEAR(p[0].depth, 0.15f, 1e-4f);
        CHECK_NEAR(p[0].normal.z, 1.0f, 1e-4f);
    }
    pxg_shape cap = pxg::no_shape();
    cap.enabled = 1;
    cap.type = PXG_CAPSULE;
    cap.capsule_radius = 0.2f;
    cap.capsule_half_height = 0.5f;
    cap.capsule_axis = 0;
    p = world.compute_penetration(cap, pxg::Vec3{0.0f, 0.0f, -0.3f}, pxg::Quat{});
    CHECK(p.size() == 1);
    if (p.size() == 1) {
        CHECK_NEAR(p[0].depth, 0.5f, 1e-4f);
        CHECK_NEAR(p[0].normal.z, 1.0f, 1e-4f);
    }
    CHECK(world.compute_penetration(sphere, pxg::Vec3{0.0f, 0.0f, 0.6f}, pxg::Quat{}).empty());
    pxg::QueryFilter ignore_floor;
    ignore_floor.ignore = floor_id;
    CHECK(world.compute_penetration(sphere, pxg::Vec3{0.0f, 0.0f, 0.4f}, pxg::Quat{}, ignore_floor).empty());
}

// The write-back of system/physics_system.h:1058-1075: each primary dynamic body's render pose,
// glm::mix of the previous and current position and glm::slerp of the orientations at
// interpolation_alpha() (current pose when the body's interpolation is None).
struct RenderPose {
    float p[3], q[4];
};
RenderPose slerp_pose(const pxg_pose& b, float alpha) {
    RenderPose r;
    if (b.interpolation == PXG_INTERPOLATE_NONE) {
        std::memcpy(r.p, b.position, sizeof(r.p));
        std::memcpy(r.q, b.orientation, sizeof(r.q));
        return r;
    }
    for (int i = 0; i < 3; ++i) r.p[i] = b.prev_position[i] * (1.0f - alpha) + b.position[i] * alpha;
    const float* a = b.prev_orientation;
    float c[4] = {b.orientation[0], b.orientation[1], b.orientation[2], b.orientation[3]};
    float cosv = a[0] * c[0] + a[1] * c[1] + a[2] * c[2] + a[3] * c[3];
    if (cosv < 0.0f) {  // the shorter arc, as glm::slerp takes it
        for (float& x : c) x = -x;
        cosv = -cosv;
    }
    if (cosv > 1.0f - 1e-6f) {
        for (int i = 0; i < 4; ++i) r.q[i] = a[i] * (1.0f - alpha) + c[i] * alpha;
    } else {
        const float angle = std::acos(cosv);
        const float wa = std::sin((1.0f - alpha) * angle) / std::sin(angle), wc = std::sin(alpha * angle) / std::sin(angle);
        for (int i = 0; i < 4; ++i) r.q[i] = a[i] * wa + c[i] * wc;
    }
    return r;
}
pxg_pose pose_of(const pxg_body& b) {
    pxg_pose p;
    std::memcpy(p.position, b.position, sizeof(p.position));
    std::memcpy(p.orientation, b.orientation, sizeof(p.orientation));
    std::memcpy(p.prev_position, b.prev_position, sizeof(p.prev_position));
    std::memcpy(p.prev_orientation, b.prev_orientation, sizeof(p.prev_orientation));
    p.type = b.type;
    p.interpolation = b.interpolation;
    return p;
}

// The transform write-back run as the reference runs it, from job workers: four threads each
// take a range of bodies and read them through get_body() (the non-const overload, as
// PhysicsSystem does) and through pose(); both give the serial pass's render poses bit for bit.
// A read-only pass uploads nothing before the next step; a velocity edited through get_body() in a
// worker uploads that one body, and the world then steps exactly as one whose whole body table
// was re-uploaded with the same edit.
void write_back_from_worker_threads_matches_a_serial_pass() {
    auto build = [](pxg::PhysicsWorld& w, std::vector<pxg::BodyId>& ids) {
        add_static_ground(w);
        for (int i = 0; i < 96; ++i) {
            pxg_body b = body(PXG_DYNAMIC, -2.5f + 0.55f * float(i % 10), -2.5f + 0.55f * float((i / 10) % 10),
                              0.6f + 0.7f * float(i / 100) + 0.01f * float(i % 7));
            b.angular_velocity[2] = 0.3f * float(i % 5);
            if (i % 9 == 0) b.interpolation = PXG_INTERPOLATE_NONE;
            ids.push_back(w.add_body(b, box_shape(0.25f, 0.25f, 0.25f)));
        }
    };
    pxg::PhysicsWorld world, twin;
    std::vector<pxg::BodyId> ids, twin_ids;
    build(world, ids);
    build(twin, twin_ids);
    for (int i = 0; i < 20; ++i) {
        world.step_fixed(kH);
        twin.step_fixed(kH);
    }
    CHECK(world.step(2.5f * kH) == 2u);
    CHECK_NEAR(world.interpolation_alpha(), 0.5f, 1e-4f);
    twin.step_fixed(kH);
    twin.step_fixed(kH);
    const float alpha = world.interpolation_alpha();
    CHECK(world.pose(ids[0])->interpolation == PXG_INTERPOLATE_NONE && world.pose(ids[1])->interpolation == PXG_INTERPOLATE);
    // a new step: the workers' first calls download the state, the others wait for it
    world.step_fixed(kH);
    twin.step_fixed(kH);
    std::vector<RenderPose> by_body(ids.size()), by_pose(ids.size());
    auto worker = [&](size_t begin, size_t end) {
        for (size_t k = begin; k < end; ++k) {
            pxg_body* b = world.get_body(ids[k]);
            if (!b || b->type != PXG_DYNAMIC) continue;
            by_body[k] = slerp_pose(pose_of(*b), alpha);
            by_pose[k] = slerp_pose(*world.pose(ids[k]), alpha);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < 4; ++t) pool.emplace_back(worker, ids.size() * t / 4, ids.size() * (t + 1) / 4);
    for (auto& th : pool) th.join();
    for (size_t k = 0; k < ids.size(); ++k) {
        const RenderPose serial = slerp_pose(pose_of(*world.get_body(ids[k])), alpha);
        CHECK(std::memcmp(&by_body[k], &serial, sizeof(RenderPose)) == 0);
        CHECK(std::memcmp(&by_pose[k], &serial, sizeof(RenderPose)) == 0);
    }
    world.step_fixed(kH);
    twin.step_fixed(kH);
    CHECK(world.bodies_uploaded_last_step() == 0u);
    // one worker edits one body's velocity; the others only read
    std::vector<std::thread> edit;
    for (int t = 0; t < 4; ++t)
        edit.emplace_back([&, t]() {
            for (size_t k = size_t(t); k < ids.size(); k += 4) {
                pxg_body* b = world.get_body(ids[k]);
                if (k == 37) b->linear_velocity[0] += 1.5f;
            }
        });
    for (auto& th : edit) th.join();
    {   // the twin gets the same edit through a full re-upload of its table
        pxg_world* h = twin.handle();
        const uint32_t n = twin.body_count();  // the ground and the 96 boxes
        std::vector<pxg_body> all(n);
        CHECK(pxg_get_bodies(h, all.data(), n) == PXG_OK);
        all[twin_ids[37].index].linear_velocity[0] += 1.5f;
        CHECK(pxg_set_bodies(h, all.data(), n) == PXG_OK);
    }
    world.step_fixed(kH);
    CHECK(world.bodies_uploaded_last_step() == 1u);
    for (int i = 0; i < 30; ++i) {
        if (i) world.step_fixed(kH);
        pxg_world* h = twin.handle();
        CHECK(pxg_step_fixed(h, kH) == PXG_OK);
        uint64_t ha = 0, hb = 0;
        CHECK(pxg_state_hash(world.handle(), &ha) == PXG_OK && pxg_state_hash(h, &hb) == PXG_OK);
        CHECK(ha == hb);
    }
}

// The batched overlap and cast methods against the single-query ones on a small scene whose body
// slots carry later generations (a removed and re-added body): the same BodyIds in the same order,
// lists cut at the capacity while the counts stay whole, and the same cast hits.
void batched_overlaps_and_casts_equal_single_queries() {
    pxg::PhysicsWorld world;
    add_static_ground(world);
    std::vector<pxg::BodyId> ids;
    for (int i = 0; i < 40; ++i) {
        pxg_body b = body(PXG_DYNAMIC, -1.5f + 0.5f * float(i % 7), -1.5f + 0.5f * float(i / 7), 0.25f);
        ids.push_back(world.add_body(b, i % 2 ? sphere_shape(0.22f) : box_shape(0.2f, 0.2f, 0.2f)));
    }
    world.remove_body(ids[5]);
    world.remove_body(ids[12]);
    world.step_fixed(kH);
    ids[5] = world.add_body(body(PXG_DYNAMIC, 0.1f, 0.2f, 0.3f), sphere_shape(0.3f));
    ids[12] = world.add_body(body(PXG_DYNAMIC, -0.4f, 0.7f, 0.3f), box_shape(0.3f, 0.2f, 0.25f));
    CHECK(ids[5].generation == 1u && ids[12].generation == 1u);
    for (int i = 0; i < 10; ++i) world.step_fixed(kH);
    std::vector<pxg_overlap_query> qs;
    for (int i = 0; i < 24; ++i) {
        pxg_overlap_query q;
        std::memset(&q, 0, sizeof(q));
        const float x = -1.4f + 0.13f * float(i), y = 1.0f - 0.09f * float(i);
        q.type = uint32_t(i % 3);
        q.center[0] = x; q.center[1] = y; q.center[2] = 0.3f;
        q.radius = 0.3f + 0.05f * float(i % 5);
        q.half_extents[0] = 0.4f; q.half_extents[1] = 0.3f; q.half_extents[2] = 0.2f;
        q.orientation[2] = std::sin(0.1f * float(i)); q.orientation[3] = std::cos(0.1f * float(i));
        q.a[0] = x - 0.5f; q.a[1] = y; q.a[2] = 0.2f;
        q.b[0] = x + 0.5f; q.b[1] = y + 0.3f; q.b[2] = 0.3f;
        qs.push_back(q);
    }
    for (uint32_t cap : {64u, 2u}) {
        std::vector<std::vector<pxg::BodyId>> lists;
        std::vector<uint32_t> counts;
        world.overlap_batch(qs, lists, counts, cap);
        bool truncated = false;
        for (size_t i = 0; i < qs.size(); ++i) {
            const pxg_overlap_query& q = qs[i];
            std::vector<pxg::BodyId> one;
            if (q.type == PXG_SPHERE) one = world.overlap_sphere(pxg::Vec3{q.center[0], q.center[1], q.center[2]}, q.radius);
            else if (q.type == PXG_BOX)
                one = world.overlap_box(pxg::Vec3{q.center[0], q.center[1], q.center[2]},
                                        pxg::Vec3{q.half_extents[0], q.half_extents[1], q.half_extents[2]},
                                        pxg::Quat{q.orientation[0], q.orientation[1], q.orientation[2], q.orientation[3]});
            else one = world.overlap_capsule(pxg::Vec3{q.a[0], q.a[1], q.a[2]}, pxg::Vec3{q.b[0], q.b[1], q.b[2]}, q.radius);
            CHECK(counts[i] == one.size());
            CHECK(lists[i].size() == std::min<size_t>(cap, one.size()));
            for (size_t k = 0; k < lists[i].size(); ++k) CHECK(lists[i][k] == one[k] && world.is_valid(lists[i][k]));
            truncated = truncated || one.size() > cap;
        }
        if (cap == 2u) CHECK(truncated);
    }
    std::vector<pxg_cast_query> casts;
    for (int i = 0; i < 18; ++i) {
        pxg_cast_query c;
        std::memset(&c, 0, sizeof(c));
        c.shape = pxg::no_shape();
        c.shape.enabled = 1;
        c.shape.type = uint32_t(i % 3);
        c.shape.radius = 0.15f;
        c.shape.half_extents[0] = 0.1f; c.shape.half_extents[1] = 0.15f; c.shape.half_extents[2] = 0.1f;
        c.shape.capsule_radius = 0.1f;
        c.shape.capsule_half_height = 0.2f;
        c.shape.capsule_axis = 0;
        c.position[0] = -1.6f + 0.18f * float(i); c.position[1] = -1.2f + 0.1f * float(i); c.position[2] = 2.0f;
        c.orientation[3] = 1.0f;
        c.direction[2] = -1.0f;
        c.max_distance = 5.0f;
        casts.push_back(c);
    }
    std::vector<pxg::RaycastHit> hits;
    std::vector<uint8_t> found;
    world.shape_cast_batch(casts, hits, found);
    int n_found = 0;
    for (size_t i = 0; i < casts.size(); ++i) {
        const pxg_cast_query& c = casts[i];
        const pxg::Vec3 o{c.position[0], c.position[1], c.position[2]}, d{0.0f, 0.0f, -1.0f};
        pxg::RaycastHit one;
        bool hit;
        if (c.shape.type == PXG_SPHERE) hit = world.sphere_cast(o, 0.15f, d, 5.0f, one);
        else if (c.shape.type == PXG_BOX) hit = world.box_cast(o, pxg::Vec3{0.1f, 0.15f, 0.1f}, pxg::Quat{}, d, 5.0f, one);
        else hit = world.capsule_cast(o, 0.1f, 0.2f, 0, pxg::Quat{}, d, 5.0f, one);
        CHECK(hit == (found[i] != 0));
        if (!hit) continue;
        ++n_found;
        CHECK(hits[i].body == one.body && hits[i].shape_index == one.shape_index && hits[i].distance == one.distance);
        CHECK(hits[i].point.x == one.point.x && hits[i].normal.z == one.normal.z);
    }
    CHECK(n_found > 9);
}

// A floor of n x n quads over [-2, 2]^2 at z = 0 (narrowphase_test.cpp:314-327), as a soup.
struct FloorSoup {
    std::vector<float> positions;
    std::vector<uint32_t> indices;
    explicit FloorSoup(int n) {
        const float step = 4.0f / float(n);
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) positions.insert(positions.end(), {-2.0f + float(i) * step, -2.0f + float(j) * step, 0.0f});
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const uint32_t a = uint32_t(j * (n + 1) + i), b = a + 1, c = a + uint32_t(n) + 2, d = a + uint32_t(n) + 1;
                indices.insert(indices.end(), {a, b, c, a, c, d});
            }
    }
    pxg_mesh_desc desc() const {
        return pxg_mesh_desc{positions.data(), uint32_t(positions.size() / 3), uint32_t(indices.size()), indices.data()};
    }
};
pxg_shape mesh_shape(int32_t index) {
    pxg_shape s = box_shape(0, 0, 0);
    s.type = PXG_TRIANGLE_MESH;
    s.capsule_axis = index;
    s.radius = 1.0f;  // scale
    return s;
}

// narrowphase_test.cpp:432-450: a box whose centre has sunk under a mesh floor is lifted back out
void embedded_box_climbs_back_onto_a_mesh_floor() {
    pxg::PhysicsWorld world;
    FloorSoup floor(8);
    const pxg_mesh_desc d = floor.desc();
    world.set_meshes(&d, 1);
    world.add_body(body(PXG_STATIC, 0, 0, 0), mesh_shape(0));
    pxg::BodyId box = add_dynamic_box(world, 0.3f, -0.4f, -0.1f, 0.11f, 0.11f, 0.11f, 1.2f);
    for (int i = 0; i < 240; ++i) world.step_fixed(kH);
    const pxg_body* b = world.get_body(box);
    CHECK_NEAR(b->position[2], 0.11, 0.02);
    CHECK(vlen(b->linear_velocity) < 0.05f);
}

// A cube hull built by hand (geometry::ConvexHull's arrays) dropped on the mesh floor rests at its
// half height; the tables survive the world being re-created to grow past its body capacity.
void cube_hull_rests_on_a_mesh_floor_across_world_growth() {
    pxg_config cfg = pxg::default_config();
    cfg.max_bodies = cfg.max_shapes = 8;
    pxg::PhysicsWorld world(cfg);
    const float h = 0.3f;
    std::vector<float> v;
    for (int k = 0; k < 8; ++k) v.insert(v.end(), {(k & 1) ? h : -h, (k & 2) ? h : -h, (k & 4) ? h : -h});
    // faces counter-clockwise about their outward normals: -x, +x, -y, +y, -z, +z
    const std::vector<uint32_t> fv = {0, 4, 6, 2, 1, 3, 7, 5, 0, 1, 5, 4, 2, 6, 7, 3, 0, 2, 3, 1, 4, 5, 7, 6};
    const std::vector<uint32_t> starts = {0, 4, 8, 12, 16, 20, 24};
    const std::vector<float> planes = {-1, 0, 0, h, 1, 0, 0, h, 0, -1, 0, h, 0, 1, 0, h, 0, 0, -1, h, 0, 0, 1, h};
    std::vector<uint32_t> edges;
    for (uint32_t a = 0; a < 8; ++a)
        for (uint32_t bit = 1; bit < 8; bit <<= 1)
            if (!(a & bit)) edges.insert(edges.end(), {a, a | bit});
    pxg_hull hull;
    std::memset(&hull, 0, sizeof(hull));
    hull.vertices = v.data(); hull.n_vertices = 8; hull.n_faces = 6;
    hull.face_planes = planes.data(); hull.face_starts = starts.data(); hull.face_vertices = fv.data();
    hull.edges = edges.data(); hull.n_edges = uint32_t(edges.size() / 2);
    for (int i = 0; i < 3; ++i) { hull.local_min[i] = -h; hull.local_max[i] = h; }
    hull.volume = 8.0f * h * h * h;
    world.set_hulls(&hull, 1);
    FloorSoup floor(16);
    const pxg_mesh_desc d = floor.desc();
    world.set_meshes(&d, 1);
    world.add_body(body(PXG_STATIC, 0, 0, 0), mesh_shape(0));
    pxg_body hb = body(PXG_DYNAMIC, 0.2f, 0.1f, 1.0f);
    const float inv = 1.0f / (1.0f / 3.0f * (h * h + h * h));
    hb.inv_inertia_local[0] = hb.inv_inertia_local[1] = hb.inv_inertia_local[2] = inv;
    pxg_shape hs = box_shape(0, 0, 0);
    hs.type = PXG_CONVEX_HULL;
    hs.capsule_axis = 0;
    hs.radius = 1.0f;
    pxg::BodyId cube = world.add_body(hb, hs);
    for (int i = 0; i < 90; ++i) world.step_fixed(kH);
    for (int i = 0; i < 12; ++i) add_dynamic_sphere(world, -1.5f + 0.25f * float(i), 1.5f, 0.5f, 0.1f);  // grows the world
    for (int i = 0; i < 150; ++i) world.step_fixed(kH);
    const pxg_body* b = world.get_body(cube);
    CHECK_NEAR(b->position[2], h, 0.02);
    CHECK(best_up_alignment(b->orientation) > 0.999f);
    CHECK(vlen(b->linear_velocity) < 0.05f);
}

// cloth_solver_test.cpp:81-371 through the facade: a sheet pinned to a kinematic ball follows it
// across uneven frames (some with no fixed step), the particles outside the pins drape over the
// ball without entering it, and the cloth is part of the state hash; a removed cloth's handle is
// stale and its slot is reused under a new generation.
void cloth_pinned_to_a_moving_ball_follows_it_and_drapes() {
    pxg::PhysicsWorld world;
    pxg_body kb = body(PXG_KINEMATIC, 0, 0, 3);
    kb.use_gravity = 0;
    kb.mass = 1.0f;
    pxg::BodyId ball = world.add_body(kb, sphere_shape(1.0f));
    pxg::GridClothDesc d;
    d.columns = d.rows = 15;
    d.width = d.height = 4.0f;
    d.center[2] = 4.05f;
    pxg::Cloth c = pxg::make_grid_cloth(d);
    const float pin[3] = {0.0f, 0.0f, 4.0f};
    const uint32_t pinned = pxg::pin_to_body(c, ball.index, ball.generation, *world.get_body(ball), pin, 0.4f);
    pxg::build_tethers(c);
    CHECK(pinned > 0 && !c.tethers.empty());
    pxg::ClothId id = world.add_cloth(std::move(c));
    CHECK(world.is_valid(id) && world.cloth_count() == 1);
    for (int i = 0; i < 120; ++i) world.step(kH);
    const uint64_t h1 = world.world_state_hash();
    std::vector<float> rest;
    {
        const pxg::Cloth* cl = world.get_cloth(id);
        const float* bp = world.get_body(ball)->position;
        for (const pxg_cloth_anchor& a : cl->anchors)
            for (int k = 0; k < 3; ++k) rest.push_back(cl->particles[a.particle].position[k] - bp[k]);
    }
    for (int f = 0; f < 120; ++f) {
        const float dt = f % 2 ? kH * 1.1f : kH * 0.9f;
        pxg_body* b = world.get_body(ball);
        b->position[0] += 2.0f * dt;
        b->linear_velocity[0] = 2.0f;
        world.step(dt);
    }
    const pxg::Cloth* cl = world.get_cloth(id);
    const float* bp = world.get_body(ball)->position;
    CHECK(bp[0] > 0.4f);
    for (size_t k = 0; k < cl->anchors.size(); ++k)
        for (int j = 0; j < 3; ++j) CHECK_NEAR(cl->particles[cl->anchors[k].particle].position[j] - bp[j], rest[3 * k + j], 2e-3);
    for (const pxg_particle& p : cl->particles) {
        const float dx = p.position[0] - bp[0], dy = p.position[1] - bp[1], dz = p.position[2] - bp[2];
        CHECK(std::sqrt(dx * dx + dy * dy + dz * dz) > 0.98f);
    }
    CHECK(h1 != 0);
    world.remove_cloth(id);
    CHECK(!world.is_valid(id) && world.get_cloth(id) == nullptr && world.cloth_count() == 0);
    pxg::ClothId again = world.add_cloth(pxg::make_grid_cloth(d));
    CHECK(again.index == id.index && again.generation == id.generation + 1 && !world.is_valid(id));
    world.step(kH);
    CHECK(world.get_cloth(again)->particles[0].position[2] < 4.05f);
}

// The same pinned drape run twice gives the same state hash, and the hash covers the cloth.
void cloth_drape_is_deterministic_and_hashed() {
    auto run = [](bool with_cloth) {
        pxg::PhysicsWorld world;
        world.add_body(body(PXG_STATIC, 0, 0, 3), sphere_shape(1.0f));
        if (with_cloth) {
            pxg::GridClothDesc d;
            d.columns = d.rows = 13;
            d.width = d.height = 4.0f;
            d.center[2] = 4.05f;
            d.params.self_collision = 1;
            d.params.wind[0] = 0.4f;
            d.params.wind_turbulence = 0.6f;
            pxg::Cloth c = pxg::make_grid_cloth(d);
            const float pin[3] = {0.0f, 0.0f, 4.05f};
            pxg::pin_static(c, pin, 0.4f);
            pxg::build_tethers(c);
            world.add_cloth(std::move(c));
        }
        for (int i = 0; i < 90; ++i) world.step(kH);
        return world.world_state_hash();
    };
    const uint64_t a = run(true);
    CHECK(a == run(true));
    CHECK(a != run(false));
}

// hair_solver_test.cpp through the facade: roots stay pinned, a moving root carries its strand
// without stretching it, and a groom colliding only with listed bodies ignores the others.
void strands_follow_their_roots_and_collide_only_with_listed_bodies() {
    pxg::PhysicsWorld world;
    // two horizontal strands that fall and hang through where a sphere sits under each root
    pxg::BodyId listed = world.add_body(body(PXG_STATIC, -2.0f, 0, -0.35f), sphere_shape(0.3f));
    world.add_body(body(PXG_STATIC, 2.0f, 0, -0.35f), sphere_shape(0.3f));
    std::vector<pxg::StrandDesc> descs(2);
    for (int k = 0; k < 8; ++k) descs[0].points.insert(descs[0].points.end(), {-2.2f + 0.05f * float(k), 0.0f, 0.0f});
    for (int k = 0; k < 8; ++k) descs[1].points.insert(descs[1].points.end(), {1.8f + 0.05f * float(k), 0.0f, 0.0f});
    pxg_strand_params sp = pxg::default_strand_params();
    sp.global_shape = sp.local_shape = 0.0f;
    pxg::StrandSet set = pxg::make_strands(descs, sp);
    CHECK(set.particles.size() == 16 && set.stretch.size() == 14 && set.stretch_batches.size() == 4);
    set.collide_bodies.push_back(world.body_ref(listed));
    pxg::StrandSetId id = world.add_strands(std::move(set));
    CHECK(world.is_valid(id) && world.strand_set_count() == 1);
    for (int i = 0; i < 180; ++i) world.step(kH);
    const pxg::StrandSet* s = world.get_strands(id);
    auto dist = [](const float* p, float cx) {
        const float dx = p[0] - cx, dz = p[2] + 0.35f;
        return std::sqrt(dx * dx + p[1] * p[1] + dz * dz);
    };
    float near_listed = 1e9f, near_other = 1e9f;
    for (int k = 0; k < 8; ++k) near_listed = std::min(near_listed, dist(s->particles[k].position, -2.0f));
    for (int k = 8; k < 16; ++k) near_other = std::min(near_other, dist(s->particles[k].position, 2.0f));
    CHECK(near_listed > 0.3f);   // kept out of the listed sphere
    CHECK(near_other < 0.25f);   // hangs through the sphere it does not collide with
    CHECK(s->particles[0].position[2] == 0.0f && s->particles[8].position[2] == 0.0f);
    for (int f = 0; f < 60; ++f) {
        pxg::StrandSet* m = world.get_strands(id);
        pxg::set_strand_roots(*m, {-2.2f + 0.01f * float(f + 1), 0.0f, 0.0f, 1.8f, 0.0f, 0.0f});
        world.step(kH);
    }
    s = world.get_strands(id);
    CHECK_NEAR(s->particles[0].position[0], -1.6, 1e-5);
    float len = 0.0f;
    for (int k = 0; k + 1 < 8; ++k) {
        const float* a = s->particles[k].position;
        const float* b = s->particles[k + 1].position;
        len += std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
    }
    CHECK_NEAR(len, 0.35, 0.01);
    world.remove_strands(id);
    CHECK(!world.is_valid(id) && world.get_strands(id) == nullptr);
}

int main() {
    struct T {
        const char* name;
        void (*fn)();
    };
    const T tests[] = {
        {"free_fall_matches_analytic_while_static_and_kinematic_hold", free_fall_matches_analytic_while_static_and_kinematic_hold},
        {"stale_body_handle_is_rejected_after_slot_reuse", stale_body_handle_is_rejected_after_slot_reuse},
        {"resting_sphere_falls_asleep_at_rest_height", resting_sphere_falls_asleep_at_rest_height},
        {"moving_kinematic_platform_wakes_sleeping_body", moving_kinematic_platform_wakes_sleeping_body},
        {"spinning_kinematic_body_wakes_sleeping_body_it_sweeps", spinning_kinematic_body_wakes_sleeping_body_it_sweeps},
        {"sleeping_body_ignores_awake_neighbours_impulse", sleeping_body_ignores_awake_neighbours_impulse},
        {"restitution_one_bounces_high_and_zero_never_bounces", restitution_one_bounces_high_and_zero_never_bounces},
        {"fast_box_does_not_tunnel_through_thin_wall", fast_box_does_not_tunnel_through_thin_wall},
        {"ten_box_stack_settles_flat_and_sleeps", ten_box_stack_settles_flat_and_sleeps},
        {"mass_ratio_100_to_1_stack_stays_stable", mass_ratio_100_to_1_stack_stays_stable},
        {"same_scene_twice_hashes_the_same", same_scene_twice_hashes_the_same},
        {"on_substep_slot_changes_velocity_and_defers_creation", on_substep_slot_changes_velocity_and_defers_creation},
        {"hinge_door_swings_and_stops_at_its_limit", hinge_door_swings_and_stops_at_its_limit},
        {"hinge_free_pendulum_swings_without_drift", hinge_free_pendulum_swings_without_drift},
        {"hinge_zero_range_limit_acts_rigid", hinge_zero_range_limit_acts_rigid},
        {"ball_joint_holds_anchor_and_releases_on_remove", ball_joint_holds_anchor_and_releases_on_remove},
        {"cone_twist_respects_swing_and_twist_limits", cone_twist_respects_swing_and_twist_limits},
        {"joint_connected_bodies_share_a_sleep_island", joint_connected_bodies_share_a_sleep_island},
        {"removing_a_body_invalidates_its_joints_and_stale_joint_handles_are_rejected",
         removing_a_body_invalidates_its_joints_and_stale_joint_handles_are_rejected},
        {"pair_collision_switch_lets_a_sphere_fall_through_the_ground", pair_collision_switch_lets_a_sphere_fall_through_the_ground},
        {"trigger_fires_enter_and_exit_once_with_stays_between", trigger_fires_enter_and_exit_once_with_stays_between},
        {"bouncing_sphere_enters_and_exits_and_step_clears_the_queue",
         bouncing_sphere_enters_and_exits_and_step_clears_the_queue},
        {"raycasts_hit_boxes_with_exact_distances_and_filters", raycasts_hit_boxes_with_exact_distances_and_filters},
        {"overlaps_find_only_touched_bodies", overlaps_find_only_touched_bodies},
        {"sweeps_are_exact_against_box_corner_and_capsule", sweeps_are_exact_against_box_corner_and_capsule},
        {"compute_penetration_is_exact_against_box", compute_penetration_is_exact_against_box},
        {"embedded_box_climbs_back_onto_a_mesh_floor", embedded_box_climbs_back_onto_a_mesh_floor},
        {"cube_hull_rests_on_a_mesh_floor_across_world_growth", cube_hull_rests_on_a_mesh_floor_across_world_growth},
        {"write_back_from_worker_threads_matches_a_serial_pass", write_back_from_worker_threads_matches_a_serial_pass},
        {"batched_overlaps_and_casts_equal_single_queries", batched_overlaps_and_casts_equal_single_queries},
        {"cloth_pinned_to_a_moving_ball_follows_it_and_drapes", cloth_pinned_to_a_moving_ball_follows_it_and_drapes},
        {"cloth_drape_is_deterministic_and_hashed", cloth_drape_is_deterministic_and_hashed},
        {"strands_follow_their_roots_and_collide_only_with_listed_bodies",
         strands_follow_their_roots_and_collide_only_with_listed_bodies},
    };
    for (const T& t : tests) {
        int before = g_failures;
        try {
            t.fn();
        } catch (const std::exception& e) {
            std::printf("  FAILED: exception: %s\n", e.what());
            ++g_failures;
        }
        std::printf("%s %s\n", g_failures == before ? "PASS" : "FAIL", t.name);
        std::fflush(stdout);
    }
    std::printf("%d failure(s)\n", g_failures);
    return g_failures ? 1 : 0;
}
