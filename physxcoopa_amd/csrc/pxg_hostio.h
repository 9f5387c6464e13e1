// Host-side packing between the ABI records (pxg_body, pxg_shape) and the device SoA layout
// shared by the single-world and batched paths.
#pragma once
#include "../../include/pxg.h"
#include "pxg_state.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

namespace pxg {

struct BodySoA {
    std::vector<float4> pos, rot, vel, angv, force, torque, iil, ppos, prot;
    std::vector<uint32_t> flags;
    void resize(size_t n) {
        pos.resize(n); rot.resize(n); vel.resize(n); angv.resize(n); force.resize(n); torque.resize(n);
        iil.resize(n); ppos.resize(n); prot.resize(n); flags.resize(n);
    }
};

// returns false on a bad body type
inline bool pack_bodies(const pxg_body* in, size_t n, BodySoA& o) {
    o.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const pxg_body& b = in[i];
        if (b.type > 2) return false;
        o.pos[i] = make_float4(b.position[0], b.position[1], b.position[2], b.sleep_timer);
        o.rot[i] = make_float4(b.orientation[0], b.orientation[1], b.orientation[2], b.orientation[3]);
        o.vel[i] = make_float4(b.linear_velocity[0], b.linear_velocity[1], b.linear_velocity[2], b.inv_mass);
        o.angv[i] = make_float4(b.angular_velocity[0], b.angular_velocity[1], b.angular_velocity[2], b.mass);
        o.force[i] = make_float4(b.force[0], b.force[1], b.force[2], b.linear_drag);
        o.torque[i] = make_float4(b.torque[0], b.torque[1], b.torque[2], b.angular_drag);
        o.iil[i] = make_float4(b.inv_inertia_local[0], b.inv_inertia_local[1], b.inv_inertia_local[2], 0.0f);
        o.ppos[i] = make_float4(b.prev_position[0], b.prev_position[1], b.prev_position[2], 0.0f);
        o.prot[i] = make_float4(b.prev_orientation[0], b.prev_orientation[1], b.prev_orientation[2], b.prev_orientation[3]);
        o.flags[i] = (b.type & BF_TYPE) | (b.alive ? BF_ALIVE : 0u) | (b.awake ? BF_AWAKE : 0u) |
                     (b.use_gravity ? BF_GRAVITY : 0u) | ((b.constraints & 63u) << BF_CONSTRAINT_SHIFT);
    }
    return true;
}

inline void unpack_bodies(const BodySoA& s, const std::vector<float4>& iiw, size_t n, pxg_body* out) {
    for (size_t i = 0; i < n; ++i) {
        pxg_body& b = out[i];
        std::memset(&b, 0, sizeof(b));
        uint32_t f = s.flags[i];
        b.type = f & BF_TYPE;
        b.alive = (f & BF_ALIVE) ? 1u : 0u;
        b.awake = (f & BF_AWAKE) ? 1u : 0u;
        b.use_gravity = (f & BF_GRAVITY) ? 1u : 0u;
        b.constraints = (f >> BF_CONSTRAINT_SHIFT) & 63u;
        const float4 &p = s.pos[i], &r = s.rot[i], &v = s.vel[i], &w = s.angv[i], &fo = s.force[i], &to = s.torque[i];
        b.position[0] = p.x; b.position[1] = p.y; b.position[2] = p.z; b.sleep_timer = p.w;
        b.orientation[0] = r.x; b.orientation[1] = r.y; b.orientation[2] = r.z; b.orientation[3] = r.w;
        b.linear_velocity[0] = v.x; b.linear_velocity[1] = v.y; b.linear_velocity[2] = v.z; b.inv_mass = v.w;
        b.angular_velocity[0] = w.x; b.angular_velocity[1] = w.y; b.angular_velocity[2] = w.z; b.mass = w.w;
        b.force[0] = fo.x; b.force[1] = fo.y; b.force[2] = fo.z; b.linear_drag = fo.w;
        b.torque[0] = to.x; b.torque[1] = to.y; b.torque[2] = to.z; b.angular_drag = to.w;
        b.inv_inertia_local[0] = s.iil[i].x; b.inv_inertia_local[1] = s.iil[i].y; b.inv_inertia_local[2] = s.iil[i].z;
        const float* m = reinterpret_cast<const float*>(&iiw[3 * i]);
        for (int k = 0; k < 9; ++k) b.inv_inertia_world[k] = m[k];
        b.prev_position[0] = s.ppos[i].x; b.prev_position[1] = s.ppos[i].y; b.prev_position[2] = s.ppos[i].z;
        b.prev_orientation[0] = s.prot[i].x; b.prev_orientation[1] = s.prot[i].y; b.prev_orientation[2] = s.prot[i].z;
        b.prev_orientation[3] = s.prot[i].w;
    }
}

struct ShapeSoA {
    std::vector<uint4> info;
    std::vector<float4> lc, lr, dim, mat;
};
// static_owner(i) tells whether shape i's owner is a static body. Returns an error string or null.
template <class StaticFn>
inline const char* pack_shapes(const pxg_shape* in, size_t n, uint32_t n_owner_bodies, StaticFn static_owner,
                               ShapeSoA& o) {
    o.info.resize(n); o.lc.resize(n); o.lr.resize(n); o.dim.resize(n); o.mat.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const pxg_shape& s = in[i];
        if (s.owner >= n_owner_bodies) return "shape owner out of range";
        if (s.type > 2) return "only sphere, box and capsule shapes are supported";
        if (s.capsule_axis < 0 || s.capsule_axis > 2) return "bad capsule axis";
        uint32_t flags = s.type | (s.enabled ? SF_ENABLED : 0u) | (s.is_trigger ? SF_TRIGGER : 0u) |
                         (static_owner(i) ? SF_STATIC : 0u) | (uint32_t(s.capsule_axis) << 8);
        o.info[i] = make_uint4(s.owner, flags, s.layer & 31u, (s.friction_combine & 255u) | ((s.restitution_combine & 255u) << 8));
        o.lc[i] = make_float4(s.local_center[0], s.local_center[1], s.local_center[2], s.radius);
        o.lr[i] = make_float4(s.local_rotation[0], s.local_rotation[1], s.local_rotation[2], s.local_rotation[3]);
        o.dim[i] = make_float4(s.half_extents[0], s.half_extents[1], s.half_extents[2], s.capsule_radius);
        o.mat[i] = make_float4(s.capsule_half_height, s.dynamic_friction, s.restitution, s.static_friction);
    }
    return nullptr;
}

}  // namespace pxg
