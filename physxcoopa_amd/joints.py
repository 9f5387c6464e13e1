"""Joint records for the pxg ABI: the constructors of physxcoopa's PhysicsWorld
(add_hinge_joint / add_ball_joint / add_cone_twist_joint, world.h:400-477), with the rest pose
computed the way the reference computes it (glm quaternion arithmetic in float32, no contraction).
"""
import numpy as np

from ._abi import JOINT_DTYPE, PXG_BALL, PXG_CONE_TWIST, PXG_HINGE

F = np.float32


def _q(q):
    return [F(c) for c in q]  # x, y, z, w


def quat_mul(p, q):
    """glm quaternion product p * q (x, y, z, w storage)."""
    px, py, pz, pw = _q(p)
    qx, qy, qz, qw = _q(q)
    w = pw * qw - px * qx - py * qy - pz * qz
    x = pw * qx + px * qw + py * qz - pz * qy
    y = pw * qy + py * qw + pz * qx - px * qz
    z = pw * qz + pz * qw + px * qy - py * qx
    return np.array([x, y, z, w], F)


def quat_inverse(q):
    """glm::inverse(quat): conjugate(q) / dot(q, q), dot = (x x + y y) + (z z + w w)."""
    x, y, z, w = _q(q)
    d = (x * x + y * y) + (z * z + w * w)
    return np.array([-x / d, -y / d, -z / d, w / d], F)


def _cross(a, b):
    return np.array([a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]], F)


def quat_rotate(q, v):
    """glm quat * vec3: v + ((uv * w) + uuv) * 2 with uv = cross(q.xyz, v), uuv = cross(q.xyz, uv)."""
    x, y, z, w = _q(q)
    qv = np.array([x, y, z], F)
    v = np.asarray(v, F)
    uv = _cross(qv, v)
    uuv = _cross(qv, uv)
    return (v + ((uv * w) + uuv) * F(2)).astype(F)


def normalize(v):
    v = np.asarray(v, F)
    d = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
    return (v * (F(1) / np.sqrt(d, dtype=F))).astype(F)


def _base(jtype, bodies, a, b, anchor_a, anchor_b, collide_connected):
    j = np.zeros((), JOINT_DTYPE)
    j["type"] = jtype
    j["body_a"] = a
    j["body_b"] = b
    j["valid"] = 1
    j["enabled"] = 1
    j["collide_connected"] = 1 if collide_connected else 0
    j["local_anchor_a"] = anchor_a
    j["local_anchor_b"] = anchor_b
    j["local_axis_a"] = (0, 0, 1)
    j["local_axis_b"] = (0, 0, 1)
    j["swing_limit"] = -1.0
    j["twist_min"] = 1.0
    j["twist_max"] = -1.0
    j["rest_relative_rotation"] = quat_mul(quat_inverse(bodies[a]["orientation"]), bodies[b]["orientation"])
    return j


def hinge(bodies, a, b, anchor_a, anchor_b, axis_a, axis_b, use_limits=False, min_angle=0.0, max_angle=0.0,
          collide_connected=False):
    """add_hinge_joint (world.h:400-422)."""
    j = _base(PXG_HINGE, bodies, a, b, anchor_a, anchor_b, collide_connected)
    j["local_axis_a"] = axis_a
    j["local_axis_b"] = axis_b
    j["use_limits"] = 1 if use_limits else 0
    j["min_angle"] = min_angle
    j["max_angle"] = max_angle
    return j


def ball(bodies, a, b, anchor_a, anchor_b, collide_connected=False):
    """add_ball_joint (world.h:429-443)."""
    return _base(PXG_BALL, bodies, a, b, anchor_a, anchor_b, collide_connected)


def cone_twist(bodies, a, b, anchor_a, anchor_b, twist_axis_a, swing_limit, twist_min, twist_max,
               collide_connected=False):
    """add_cone_twist_joint (world.h:456-477): b's twist axis is a's carried through the rest pose."""
    j = _base(PXG_CONE_TWIST, bodies, a, b, anchor_a, anchor_b, collide_connected)
    ax = normalize(twist_axis_a)
    j["local_axis_a"] = ax
    j["local_axis_b"] = quat_rotate(quat_inverse(j["rest_relative_rotation"]), ax)
    j["swing_limit"] = swing_limit
    j["twist_min"] = twist_min
    j["twist_max"] = twist_max
    return j


def table(records):
    """A contiguous joint table from records (an empty table when there are none)."""
    if not records:
        return np.zeros(0, JOINT_DTYPE)
    return np.ascontiguousarray(np.array(records, dtype=JOINT_DTYPE))
