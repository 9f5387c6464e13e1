// Joint rows for gfx950: hinge, ball and cone-twist (physxcoopa/dynamics/solver.h:330-761).
//
// The functions work on register copies of the two bodies (JBody); the caller loads them from
// whichever body arrays its pass uses and stores them back, so the single-world kernels and the
// batched-scene kernel share the arithmetic. Evaluation orders follow the reference's glm code
// (see pxg_device.h); the 3x3 inverse and determinant are glm's cofactor expansions.
//
// atan2 is evaluated in double and rounded once (correctly rounded up to double's own error);
// the reference's std::atan2 on floats can differ from it by one ulp.
#pragma once
#include "pxg_device.h"

namespace pxd {

enum { J_HINGE = 0, J_BALL = 1, J_CONE_TWIST = 2 };
constexpr uint32_t JF_LIMITS = 1u, JF_SWING = 2u, JF_TWIST = 4u;

struct JBody {
    v3 x, v, w;
    quat q;
    float im;   // effective inverse mass (0 unless dynamic and awake)
    m3 I;       // effective world inverse inertia (0 unless dynamic and awake)
    bool act;   // dynamic and awake
};
struct JDef {
    uint32_t type, flags;
    v3 anchor_a, anchor_b, axis_a, axis_b;
    float min_angle, max_angle, swing_limit, twist_min, twist_max;
    quat rest;
};
struct JAcc {
    v3 point;
    float axis0, axis1, limit, swing;
};

// Device records: 6 float4 of definition, 2 float4 of accumulators per joint.
//   def[0] anchor_a, min_angle   def[1] anchor_b, max_angle   def[2] axis_a, swing_limit
//   def[3] axis_b, twist_min     def[4] rest (x y z w)        def[5] twist_max, type, flags, -
//   acc[0] point impulse, limit impulse                       acc[1] axis impulse x, y, swing impulse, -
PXD JDef load_jdef(const float4* def, uint32_t j) {
    JDef r;
    float4 a = def[6 * j], b = def[6 * j + 1], c = def[6 * j + 2], e = def[6 * j + 3], q = def[6 * j + 4],
           t = def[6 * j + 5];
    r.anchor_a = ld3(a); r.min_angle = a.w;
    r.anchor_b = ld3(b); r.max_angle = b.w;
    r.axis_a = ld3(c); r.swing_limit = c.w;
    r.axis_b = ld3(e); r.twist_min = e.w;
    r.rest = ldq(q);
    r.twist_max = t.x;
    r.type = __float_as_uint(t.y);
    r.flags = __float_as_uint(t.z);
    return r;
}
PXD JAcc load_jacc(const float4* acc, uint32_t j) {
    JAcc r;
    float4 a = ld_dev(acc, 2 * j), b = ld_dev(acc, 2 * j + 1);
    r.point = ld3(a); r.limit = a.w;
    r.axis0 = b.x; r.axis1 = b.y; r.swing = b.z;
    return r;
}
PXD void store_jacc(float4* acc, uint32_t j, const JAcc& r) {
    acc[2 * j] = st3(r.point, r.limit);
    acc[2 * j + 1] = make_float4(r.axis0, r.axis1, r.swing, 0.0f);
}

PXD float atan2_rounded(float y, float x) { return static_cast<float>(atan2(static_cast<double>(y), static_cast<double>(x))); }

// ---- 3x3 and quaternion helpers in glm's order ----
PXD m3 mmul(const m3& a, const m3& b) {  // R[c][r] = (A[0][r] B[c][0] + A[1][r] B[c][1]) + A[2][r] B[c][2]
    m3 r;
    r.c0 = mul(a, b.c0); r.c1 = mul(a, b.c1); r.c2 = mul(a, b.c2);
    return r;
}
PXD m3 msub(const m3& a, const m3& b) {
    m3 r;
    r.c0 = a.c0 - b.c0; r.c1 = a.c1 - b.c1; r.c2 = a.c2 - b.c2;
    return r;
}
PXD m3 skew(v3 v) {  // skew(v) * x == cross(v, x)
    m3 m;
    m.c0 = mk(0.0f, v.z, -v.y);
    m.c1 = mk(-v.z, 0.0f, v.x);
    m.c2 = mk(v.y, -v.x, 0.0f);
    return m;
}
PXD m3 diag3(float s) {
    m3 m;
    m.c0 = mk(s, 0.0f, 0.0f); m.c1 = mk(0.0f, s, 0.0f); m.c2 = mk(0.0f, 0.0f, s);
    return m;
}
PXD m3 zero3() { return diag3(0.0f); }
PXD float mdet(const m3& m) {  // glm compute_determinant (mat3)
    return m.c0.x * (m.c1.y * m.c2.z - m.c2.y * m.c1.z)
         - m.c1.x * (m.c0.y * m.c2.z - m.c2.y * m.c0.z)
         + m.c2.x * (m.c0.y * m.c1.z - m.c1.y * m.c0.z);
}
PXD m3 minverse(const m3& m) {  // glm compute_inverse (mat3)
    float inv = 1.0f / (m.c0.x * (m.c1.y * m.c2.z - m.c2.y * m.c1.z)
                      - m.c1.x * (m.c0.y * m.c2.z - m.c2.y * m.c0.z)
                      + m.c2.x * (m.c0.y * m.c1.z - m.c1.y * m.c0.z));
    m3 r;
    r.c0.x = (m.c1.y * m.c2.z - m.c2.y * m.c1.z) * inv;
    r.c1.x = -(m.c1.x * m.c2.z - m.c2.x * m.c1.z) * inv;
    r.c2.x = (m.c1.x * m.c2.y - m.c2.x * m.c1.y) * inv;
    r.c0.y = -(m.c0.y * m.c2.z - m.c2.y * m.c0.z) * inv;
    r.c1.y = (m.c0.x * m.c2.z - m.c2.x * m.c0.z) * inv;
    r.c2.y = -(m.c0.x * m.c2.y - m.c2.x * m.c0.y) * inv;
    r.c0.z = (m.c0.y * m.c1.z - m.c1.y * m.c0.z) * inv;
    r.c1.z = -(m.c0.x * m.c1.z - m.c1.x * m.c0.z) * inv;
    r.c2.z = (m.c0.x * m.c1.y - m.c1.x * m.c0.y) * inv;
    return r;
}
PXD quat qinverse(quat q) {  // glm::inverse(quat): conjugate / dot
    float d = (q.x * q.x + q.y * q.y) + (q.z * q.z + q.w * q.w);
    quat r;
    r.w = q.w / d; r.x = -q.x / d; r.y = -q.y / d; r.z = -q.z / d;
    return r;
}
// normalize(q + (0.5 * quat(0, dr)) * q) (solver.h:466-473, 722-731)
PXD quat rotate_by(quat q, v3 dr) {
    quat h;
    h.w = 0.0f * 0.5f; h.x = dr.x * 0.5f; h.y = dr.y * 0.5f; h.z = dr.z * 0.5f;
    quat p = qmul(h, q);
    quat s;
    s.w = q.w + p.w; s.x = q.x + p.x; s.y = q.y + p.y; s.z = q.z + p.z;
    return qnormalize(s);
}

// point_block_mass (solver.h:440-450)
PXD bool point_block_mass(v3 ra, v3 rb, float ima, float imb, const m3& Ia, const m3& Ib, m3& k) {
    m3 sa = skew(ra), sb = skew(rb);
    k = msub(msub(diag3(ima + imb), mmul(mmul(sa, Ia), sa)), mmul(mmul(sb, Ib), sb));
    return fabsf(mdet(k)) > 1e-12f;
}
// hinge_current_angle (solver.h:357-368)
PXD float hinge_angle(quat qa, quat qb, v3 axis_a, quat rest) {
    quat rel = qmul(qinverse(qa), qb);
    quat delta = qmul(rel, qinverse(rest));
    v3 axis = normalize(axis_a);
    v3 dv = mk(delta.x, delta.y, delta.z);
    float w = sclamp(delta.w, -1.0f, 1.0f);
    if (w < 0.0f) { w = -w; dv = -dv; }
    return 2.0f * atan2_rounded(dot(dv, axis), w);
}
// cone_swing (solver.h:381-393)
PXD float cone_swing(v3 ta, v3 tb, v3& swing_axis) {
    v3 c = cross(ta, tb);
    float s = len(c);
    float angle = atan2_rounded(s, dot(ta, tb));
    if (s > 1e-6f) {
        swing_axis = c / s;
    } else {
        v3 p, q;
        orthonormal_basis(ta, p, q);
        swing_axis = p;
    }
    return angle;
}
PXD void nearest_limit_side(float angle, float lo, float hi, float& sign, float& c) {  // solver.h:422-427
    float c_lo = angle - lo, c_hi = hi - angle;
    if (c_lo <= c_hi) { sign = 1.0f; c = c_lo; }
    else { sign = -1.0f; c = c_hi; }
}
PXD void angular_pair(JBody& a, JBody& b, v3 J) {  // apply_angular_impulse_pair, solver.h:104-111
    if (a.act) a.w = a.w - mul(a.I, J);
    if (b.act) b.w = b.w + mul(b.I, J);
}
// solve_angular_limit_row (solver.h:406-416)
PXD void limit_row(JBody& a, JBody& b, v3 dir, float c, float inv_h, float& acc) {
    float k = dot(dir, mul(a.I, dir)) + dot(dir, mul(b.I, dir));
    if (k <= kEps) return;
    float cdot = dot(b.w - a.w, dir);
    float lambda = -(cdot + gmax(c, 0.0f) * inv_h) / k;
    float nw = gmax(acc + lambda, 0.0f);
    float delta = nw - acc;
    acc = nw;
    angular_pair(a, b, dir * delta);
}
// apply_angular_position_correction (solver.h:456-474)
PXD void angular_correction(JBody& a, JBody& b, v3 n, float angle) {
    float k = dot(n, mul(a.I, n)) + dot(n, mul(b.I, n));
    if (k <= kEps) return;
    v3 corr = n * (angle / k);
    if (a.act) a.q = rotate_by(a.q, -mul(a.I, corr));
    if (b.act) b.q = rotate_by(b.q, mul(b.I, corr));
}
PXD void cone_axes(const JBody& a, const JBody& b, const JDef& j, v3& ta, v3& tb) {  // solver.h:371-374
    ta = normalize(qrot(a.q, j.axis_a));
    tb = normalize(qrot(b.q, j.axis_b));
}

// warm_start_joints for one live joint (solver.h:483-516); the caller has zeroed the limit
// accumulators
PXD void joint_warm(const JDef& j, const JAcc& acc, JBody& a, JBody& b) {
    if (!a.act && !b.act) return;
    v3 ra = qrot(a.q, j.anchor_a), rb = qrot(b.q, j.anchor_b);
    v3 pi = acc.point;
    if (dot(pi, pi) > 0.0f) {
        if (a.act) { a.v = a.v - pi * a.im; a.w = a.w - mul(a.I, cross(ra, pi)); }
        if (b.act) { b.v = b.v + pi * b.im; b.w = b.w + mul(b.I, cross(rb, pi)); }
    }
    if (j.type != J_HINGE) return;
    v3 wa = normalize(qrot(a.q, j.axis_a));
    v3 p, q;
    orthonormal_basis(wa, p, q);
    v3 ai = p * acc.axis0 + q * acc.axis1;
    if (dot(ai, ai) > 0.0f) angular_pair(a, b, ai);
}

// one live joint of solve_joint_velocity_pass (solver.h:548-632)
PXD void joint_velocity(const JDef& j, JAcc& acc, JBody& a, JBody& b, float inv_h) {
    if (!a.act && !b.act) return;
    if (j.type == J_HINGE) {
        v3 wa = normalize(qrot(a.q, j.axis_a));
        v3 p, q;
        orthonormal_basis(wa, p, q);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            v3 n = k == 0 ? p : q;
            float keff = dot(n, mul(a.I, n)) + dot(n, mul(b.I, n));
            if (keff <= kEps) continue;
            float cdot = dot(b.w - a.w, n);
            float lambda = -cdot / keff;
            if (k == 0) acc.axis0 += lambda;
            else acc.axis1 += lambda;
            angular_pair(a, b, n * lambda);
        }
        if (j.flags & JF_LIMITS) {
            float angle = hinge_angle(a.q, b.q, j.axis_a, j.rest);
            float sign, c;
            nearest_limit_side(angle, j.min_angle, j.max_angle, sign, c);
            limit_row(a, b, wa * sign, c, inv_h, acc.limit);
        }
    } else if (j.type == J_CONE_TWIST) {
        v3 ta, tb;
        cone_axes(a, b, j, ta, tb);
        if (j.flags & JF_SWING) {
            v3 sw;
            float swing = cone_swing(ta, tb, sw);
            limit_row(a, b, -sw, j.swing_limit - swing, inv_h, acc.swing);
        }
        if (j.flags & JF_TWIST) {
            float angle = hinge_angle(a.q, b.q, j.axis_a, j.rest);
            v3 ax = ta + tb;
            float l = len(ax);
            ax = l > 1e-6f ? ax / l : ta;
            float sign, c;
            nearest_limit_side(angle, j.twist_min, j.twist_max, sign, c);
            limit_row(a, b, ax * sign, c, inv_h, acc.limit);
        }
    }
    v3 ra = qrot(a.q, j.anchor_a), rb = qrot(b.q, j.anchor_b);
    m3 K;
    if (point_block_mass(ra, rb, a.im, b.im, a.I, b.I, K)) {
        v3 va = a.v + cross(a.w, ra), vb = b.v + cross(b.w, rb);
        v3 imp = mul(minverse(K), va - vb);
        acc.point = acc.point + imp;
        if (a.act) { a.v = a.v - imp * a.im; a.w = a.w - mul(a.I, cross(ra, imp)); }
        if (b.act) { b.v = b.v + imp * b.im; b.w = b.w + mul(b.I, cross(rb, imp)); }
    }
}

// one live joint of solve_joint_position (solver.h:652-747)
PXD void joint_position(const JDef& j, JBody& a, JBody& b, float slop, float pos_corr, float max_corr) {
    if (!a.act && !b.act) return;
    if (j.type == J_HINGE) {
        v3 wa = normalize(qrot(a.q, j.axis_a));
        v3 wb = normalize(qrot(b.q, j.axis_b));
        v3 e = cross(wa, wb);
        float el = len(e);
        if (el > slop) {
            v3 n = e / el;
            float mag = gmin((el - slop) * pos_corr, max_corr);
            angular_correction(a, b, n, -mag);
        }
        if (j.flags & JF_LIMITS) {
            float angle = hinge_angle(a.q, b.q, j.axis_a, j.rest);
            float over = angle < j.min_angle ? angle - j.min_angle : (angle > j.max_angle ? angle - j.max_angle : 0.0f);
            if (fabsf(over) > slop) {
                float mag = gmin((fabsf(over) - slop) * pos_corr, max_corr);
                v3 ax = normalize(qrot(a.q, j.axis_a));
                angular_correction(a, b, ax, over > 0.0f ? -mag : mag);
            }
        }
    } else if (j.type == J_CONE_TWIST) {
        v3 ta, tb;
        cone_axes(a, b, j, ta, tb);
        if (j.flags & JF_SWING) {
            v3 sw;
            float over = cone_swing(ta, tb, sw) - j.swing_limit;
            if (over > slop) {
                float mag = gmin((over - slop) * pos_corr, max_corr);
                angular_correction(a, b, sw, -mag);
                cone_axes(a, b, j, ta, tb);
            }
        }
        if (j.flags & JF_TWIST) {
            float angle = hinge_angle(a.q, b.q, j.axis_a, j.rest);
            float over = angle < j.twist_min ? angle - j.twist_min : (angle > j.twist_max ? angle - j.twist_max : 0.0f);
            if (fabsf(over) > slop) {
                float mag = gmin((fabsf(over) - slop) * pos_corr, max_corr);
                v3 ax = ta + tb;
                float l = len(ax);
                ax = l > 1e-6f ? ax / l : ta;
                angular_correction(a, b, ax, over > 0.0f ? -mag : mag);
            }
        }
    }
    v3 ra = qrot(a.q, j.anchor_a), rb = qrot(b.q, j.anchor_b);
    v3 err = (b.x + rb) - (a.x + ra);
    float el = len(err);
    m3 K;
    if (el > slop && point_block_mass(ra, rb, a.im, b.im, a.I, b.I, K)) {
        float mag = gmin((el - slop) * pos_corr, max_corr);
        v3 p = mul(minverse(K), (err / el) * mag);
        if (a.act) {
            a.x = a.x + p * a.im;
            a.q = rotate_by(a.q, mul(a.I, cross(ra, p)));
        }
        if (b.act) {
            b.x = b.x - p * b.im;
            b.q = rotate_by(b.q, -mul(b.I, cross(rb, p)));
        }
    }
}

}  // namespace pxd
