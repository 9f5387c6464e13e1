// The reference's PhysicsWorld tests (physxcoopa/tests/physics_world_test.cpp, sleep_test.cpp,
// contact_solver_test.cpp) written against the C++ facade (include/pxg_world.hpp), which steps on
// the GPU through the C ABI. Built in-tree by physxcoopa_amd/build.py; run by
// tests/test_gpu_cpp_facade.py. Prints one line per test and exits non-zero on any failure.
#include "pxg_world.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
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

}  // namespace

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
