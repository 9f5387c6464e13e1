// Host-side packing between the ABI records (pxg_body, pxg_shape) and the device SoA layout
// shared by the single-world and batched paths.
#pragma once
#include "../../include/pxg.h"
#include "pxg_state.h"

#include <hip/hip_runtime.h>

#include <algorithm>
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
                     (b.use_gravity ? BF_GRAVITY : 0u) | ((b.constraints & 63u) << BF_CONSTRAINT_SHIFT) |
                     (b.interpolation == PXG_INTERPOLATE_NONE ? BF_INTERP_NONE : 0u);
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
        b.interpolation = (f & BF_INTERP_NONE) ? PXG_INTERPOLATE_NONE : PXG_INTERPOLATE;
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
// static_owner(i) tells whether shape i's owner is a static body. n_hulls / n_meshes: the table
// sizes convex hull and triangle mesh shapes index (0: such shapes are refused); refs, if given,
// receives 1 + the largest hull and mesh index used. Returns an error string or null.
template <class StaticFn>
inline const char* pack_shapes(const pxg_shape* in, size_t n, uint32_t n_owner_bodies, StaticFn static_owner,
                               ShapeSoA& o, uint32_t n_hulls = 0, uint32_t n_meshes = 0, uint32_t* refs = nullptr) {
    o.info.resize(n); o.lc.resize(n); o.lr.resize(n); o.dim.resize(n); o.mat.resize(n);
    uint32_t hull_refs = 0, mesh_refs = 0;
    for (size_t i = 0; i < n; ++i) {
        const pxg_shape& s = in[i];
        if (s.owner >= n_owner_bodies) return "shape owner out of range";
        const bool hull = s.type == PXG_CONVEX_HULL, mesh = s.type == PXG_TRIANGLE_MESH;
        if (s.type > PXG_TRIANGLE_MESH) return "only sphere, box, capsule, convex hull and triangle mesh shapes are supported";
        if (hull && (s.capsule_axis < 0 || uint32_t(s.capsule_axis) >= n_hulls))
            return "a convex hull shape's index (capsule_axis) is not in the uploaded hull table";
        if (mesh && (s.capsule_axis < 0 || uint32_t(s.capsule_axis) >= n_meshes))
            return "a triangle mesh shape's index (capsule_axis) is not in the uploaded mesh table";
        if (!hull && !mesh && (s.capsule_axis < 0 || s.capsule_axis > 2)) return "bad capsule axis";
        if (hull) hull_refs = std::max(hull_refs, uint32_t(s.capsule_axis) + 1);
        if (mesh) mesh_refs = std::max(mesh_refs, uint32_t(s.capsule_axis) + 1);
        const bool indexed = hull || mesh;
        uint32_t flags = s.type | (s.enabled ? SF_ENABLED : 0u) | (s.is_trigger ? SF_TRIGGER : 0u) |
                         (static_owner(i) ? SF_STATIC : 0u) | (indexed ? 0u : uint32_t(s.capsule_axis) << 8);
        o.info[i] = make_uint4(s.owner, flags, s.layer & 31u, (s.friction_combine & 255u) | ((s.restitution_combine & 255u) << 8));
        o.lc[i] = make_float4(s.local_center[0], s.local_center[1], s.local_center[2], s.radius);
        o.lr[i] = make_float4(s.local_rotation[0], s.local_rotation[1], s.local_rotation[2], s.local_rotation[3]);
        o.dim[i] = make_float4(s.half_extents[0], s.half_extents[1], s.half_extents[2], s.capsule_radius);
        if (indexed) {  // the hull or mesh index goes where a box keeps its first half extent
            uint32_t k = uint32_t(s.capsule_axis);
            std::memcpy(&o.dim[i].x, &k, sizeof(k));
        }
        o.mat[i] = make_float4(s.capsule_half_height, s.dynamic_friction, s.restitution, s.static_friction);
    }
    if (refs) {
        refs[0] = hull_refs;
        refs[1] = mesh_refs;
    }
    return nullptr;
}


// ---- joints ----
// Layers of one joint pass visiting joints[0..n) forward (or reversed): a joint goes one layer after
// the last earlier joint that shares a dynamic body with it, so each body sees its joints in the
// pass's order. Non-dynamic bodies are never written by a joint row and order nothing. Joints
// with live[j] == 0 are left out. Returns the layer count L; order lists joint indices layer by
// layer, off holds the L + 1 layer offsets.
inline uint32_t layer_joint_list(const pxg_joint* joints, size_t n, const std::vector<uint8_t>& live,
                                 const std::vector<uint32_t>& body_types, bool rev, std::vector<uint32_t>& order,
                                 std::vector<uint32_t>& off) {
    std::vector<uint32_t> last(body_types.size(), 0), layer(n, 0);
    uint32_t L = 0;
    auto dyn = [&](uint32_t b) { return b < body_types.size() && body_types[b] == PXG_DYNAMIC; };
    for (size_t t = 0; t < n; ++t) {
        size_t j = rev ? n - 1 - t : t;
        if (!live[j]) continue;
        uint32_t a = joints[j].body_a, b = joints[j].body_b, l = 0;
        if (dyn(a)) l = std::max(l, last[a]);
        if (dyn(b)) l = std::max(l, last[b]);
        layer[j] = l;
        if (dyn(a)) last[a] = l + 1;
        if (dyn(b)) last[b] = l + 1;
        L = std::max(L, l + 1);
    }
    off.assign(L + 2, 0);
    for (size_t j = 0; j < n; ++j) if (live[j]) off[layer[j] + 1]++;
    for (uint32_t l = 0; l < L; ++l) off[l + 1] += off[l];
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    order.assign(n, 0);
    for (size_t t = 0; t < n; ++t) {
        size_t j = rev ? n - 1 - t : t;
        if (live[j]) order[cur[layer[j]]++] = uint32_t(j);
    }
    off.resize(L + 1);
    return L;
}

// Device records of one joint (layout in pxg_joints.h): 6 float4 of definition, 2 of accumulators.
inline void pack_joint(const pxg_joint& jt, float4* def, float4* acc) {
    uint32_t flags = 0;
    if (jt.type == PXG_HINGE && jt.use_limits) flags |= 1u;                      // hinge limit row
    if (jt.type == PXG_CONE_TWIST && jt.swing_limit >= 0.0f) flags |= 2u;        // has_swing_limit()
    if (jt.type == PXG_CONE_TWIST && jt.twist_min <= jt.twist_max) flags |= 4u;  // has_twist_limit()
    float tf, ff;
    uint32_t ty = jt.type;
    std::memcpy(&tf, &ty, 4);
    std::memcpy(&ff, &flags, 4);
    const float* aa = jt.local_anchor_a; const float* ab = jt.local_anchor_b;
    const float* xa = jt.local_axis_a; const float* xb = jt.local_axis_b; const float* q = jt.rest_relative_rotation;
    def[0] = make_float4(aa[0], aa[1], aa[2], jt.min_angle);
    def[1] = make_float4(ab[0], ab[1], ab[2], jt.max_angle);
    def[2] = make_float4(xa[0], xa[1], xa[2], jt.swing_limit);
    def[3] = make_float4(xb[0], xb[1], xb[2], jt.twist_min);
    def[4] = make_float4(q[0], q[1], q[2], q[3]);
    def[5] = make_float4(jt.twist_max, tf, ff, 0.0f);
    acc[0] = make_float4(jt.point_impulse[0], jt.point_impulse[1], jt.point_impulse[2], jt.limit_impulse);
    acc[1] = make_float4(jt.axis_impulse[0], jt.axis_impulse[1], jt.swing_impulse, 0.0f);
}
inline void unpack_joint_acc(const float4* acc, pxg_joint& jt) {
    jt.point_impulse[0] = acc[0].x; jt.point_impulse[1] = acc[0].y; jt.point_impulse[2] = acc[0].z;
    jt.limit_impulse = acc[0].w;
    jt.axis_impulse[0] = acc[1].x; jt.axis_impulse[1] = acc[1].y;
    jt.swing_impulse = acc[1].z;
}

}  // namespace pxg
