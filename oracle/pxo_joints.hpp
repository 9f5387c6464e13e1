// ORACLE -- test infrastructure only. Nothing in the product may include, link or run this.
//
// Joints: hinge, ball and cone-twist rows of physxcoopa/dynamics/solver.h:330-761, the joint
// record of physxcoopa/dynamics/joint.h:74-135 and the joint parts of the world step
// (world.h:1310-1314 post-integrate passes, world.h:1686-1694 the collision filter,
// solver.h:808-901 islands and jointed wake-up). Evaluation orders follow glm 0.9.9.3 as
// restated in pxo_math.hpp; the 3x3 inverse and determinant are glm's cofactor expansions
// (func_matrix.inl compute_inverse / compute_determinant for mat3).
//
// atan2: the reference calls std::atan2 on floats (glibc's atan2f, which is not correctly
// rounded). Atan2Mode::Libm restates that call; Atan2Mode::Rounded evaluates atan2 in double and
// rounds once, the correctly rounded value up to double's own error -- the form a GPU can match
// bit for bit. The two differ by at most one float ulp per call (pinned by the tests).
#pragma once
#include "pxo_math.hpp"

#include <vector>

namespace pxo {

enum JointType { J_HINGE = 0, J_BALL = 1, J_CONE_TWIST = 2 };  // joint.h:57-61
enum class Atan2Mode { Libm = 0, Rounded = 1 };

template <class T> inline T atan2_of(T y, T x, Atan2Mode mode) {
    if (mode == Atan2Mode::Rounded) return T(std::atan2(double(y), double(x)));
    return std::atan2(y, x);
}

template <class T> struct Joint {  // joint.h:74-135
    int type = J_HINGE;
    uint32_t a = 0, b = 0;
    V3<T> anchor_a, anchor_b;
    V3<T> axis_a{T(0), T(0), T(1)}, axis_b{T(0), T(0), T(1)};
    bool use_limits = false;
    T min_angle = 0, max_angle = 0;
    T swing_limit = T(-1), twist_min = T(1), twist_max = T(-1);
    Q<T> rest;
    V3<T> point_impulse;
    T axis_impulse[2] = {0, 0};
    T limit_impulse = 0, swing_impulse = 0;
    bool enabled = true, valid = true, collide_connected = false;
    bool has_swing_limit() const { return type == J_CONE_TWIST && swing_limit >= T(0); }
    bool has_twist_limit() const { return type == J_CONE_TWIST && twist_min <= twist_max; }
};

// ---- 3x3 helpers in glm's order ----
template <class T> inline M3<T> m3_sub(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
    for (int c = 0; c < 3; ++c) r.c[c] = a.c[c] - b.c[c];
    return r;
}
template <class T> inline T m3_det(const M3<T>& m) {
    return m.c[0][0] * (m.c[1][1] * m.c[2][2] - m.c[2][1] * m.c[1][2])
         - m.c[1][0] * (m.c[0][1] * m.c[2][2] - m.c[2][1] * m.c[0][2])
         + m.c[2][0] * (m.c[0][1] * m.c[1][2] - m.c[1][1] * m.c[0][2]);
}
template <class T> inline M3<T> m3_inverse(const M3<T>& m) {
    T inv = T(1) / (m.c[0][0] * (m.c[1][1] * m.c[2][2] - m.c[2][1] * m.c[1][2])
                  - m.c[1][0] * (m.c[0][1] * m.c[2][2] - m.c[2][1] * m.c[0][2])
                  + m.c[2][0] * (m.c[0][1] * m.c[1][2] - m.c[1][1] * m.c[0][2]));
    M3<T> r;
    r.c[0][0] = (m.c[1][1] * m.c[2][2] - m.c[2][1] * m.c[1][2]) * inv;
    r.c[1][0] = -(m.c[1][0] * m.c[2][2] - m.c[2][0] * m.c[1][2]) * inv;
    r.c[2][0] = (m.c[1][0] * m.c[2][1] - m.c[2][0] * m.c[1][1]) * inv;
    r.c[0][1] = -(m.c[0][1] * m.c[2][2] - m.c[2][1] * m.c[0][2]) * inv;
    r.c[1][1] = (m.c[0][0] * m.c[2][2] - m.c[2][0] * m.c[0][2]) * inv;
    r.c[2][1] = -(m.c[0][0] * m.c[2][1] - m.c[2][0] * m.c[0][1]) * inv;
    r.c[0][2] = (m.c[0][1] * m.c[1][2] - m.c[1][1] * m.c[0][2]) * inv;
    r.c[1][2] = -(m.c[0][0] * m.c[1][2] - m.c[1][0] * m.c[0][2]) * inv;
    r.c[2][2] = (m.c[0][0] * m.c[1][1] - m.c[1][0] * m.c[0][1]) * inv;
    return r;
}
template <class T> inline M3<T> skew(const V3<T>& v) {  // skew(v) * x == cross(v, x)
    M3<T> m;
    m.c[0] = V3<T>(T(0), v.z, -v.y);
    m.c[1] = V3<T>(-v.z, T(0), v.x);
    m.c[2] = V3<T>(v.y, -v.x, T(0));
    return m;
}
// glm::inverse(quat): conjugate(q) / dot(q, q)
template <class T> inline Q<T> qinverse(const Q<T>& q) {
    T d = qdot(q, q);
    return Q<T>::wxyz(q.w / d, -q.x / d, -q.y / d, -q.z / d);
}
// normalize(q + (0.5 * (0, dr)) * q): the first-order orientation update of the joint position
// corrections (solver.h:466-473, 722-731)
template <class T> inline Q<T> rotate_by(const Q<T>& q, const V3<T>& dr) {
    Q<T> dq = Q<T>::wxyz(T(0), dr.x, dr.y, dr.z);
    Q<T> h = Q<T>::wxyz(dq.w * T(0.5f), dq.x * T(0.5f), dq.y * T(0.5f), dq.z * T(0.5f));
    return qnormalize(qadd(q, qmul(h, q)));
}

// point_block_mass (solver.h:440-450)
template <class T> inline bool point_block_mass(const V3<T>& ra, const V3<T>& rb, T ima, T imb, const M3<T>& Ia,
                                                const M3<T>& Ib, M3<T>& k) {
    M3<T> sa = skew(ra), sb = skew(rb);
    k = m3_sub(m3_sub(M3<T>(ima + imb), mul(mul(sa, Ia), sa)), mul(mul(sb, Ib), sb));
    return std::abs(m3_det(k)) > T(1e-12f);
}

// hinge_current_angle (solver.h:357-368)
template <class T> inline T hinge_angle(const Q<T>& qa, const Q<T>& qb, const V3<T>& axis_a, const Q<T>& rest,
                                        Atan2Mode mode) {
    Q<T> rel = qmul(qinverse(qa), qb);
    Q<T> delta = qmul(rel, qinverse(rest));
    V3<T> axis = normalize(axis_a);
    V3<T> dv(delta.x, delta.y, delta.z);
    T w = sclamp(delta.w, T(-1), T(1));
    if (w < T(0)) { w = -w; dv = -dv; }
    return T(2) * atan2_of(dot(dv, axis), w, mode);
}

// cone_swing (solver.h:381-393)
template <class T> inline T cone_swing(const V3<T>& ta, const V3<T>& tb, V3<T>& swing_axis, Atan2Mode mode) {
    V3<T> c = cross(ta, tb);
    T s = length(c);
    T angle = atan2_of(s, dot(ta, tb), mode);
    if (s > T(1e-6f)) {
        swing_axis = c / s;
    } else {
        T sign = ta.z >= T(0) ? T(1) : T(-1);  // util::orthonormal_basis, first vector
        T a = T(-1) / (sign + ta.z);
        T b = ta.x * ta.y * a;
        swing_axis = V3<T>(T(1) + sign * ta.x * ta.x * a, sign * b, -sign * ta.x);
    }
    return angle;
}

// nearest_limit_side (solver.h:422-427)
template <class T> inline void nearest_limit_side(T angle, T lo, T hi, T& sign, T& c) {
    T c_lo = angle - lo, c_hi = hi - angle;
    if (c_lo <= c_hi) { sign = T(1); c = c_lo; }
    else { sign = T(-1); c = c_hi; }
}

// The joint passes over a body array. Body<T> comes from pxo_world.hpp via the includer.
template <class T, class BodyT> struct JointSolver {
    std::vector<BodyT>& bodies;
    std::vector<Joint<T>>& joints;
    Atan2Mode mode;

    static bool act(const BodyT& b) { return b.type == 2 && b.awake; }  // Dynamic and awake

    static void angular_pair(BodyT& a, BodyT& b, const V3<T>& J) {  // solver.h:104-111
        if (act(a)) a.w -= mul(a.iiw, J);
        if (act(b)) b.w += mul(b.iiw, J);
    }
    void cone_axes(const BodyT& a, const BodyT& b, const Joint<T>& j, V3<T>& ta, V3<T>& tb) const {  // solver.h:371-374
        ta = normalize(qrot(a.q, j.axis_a));
        tb = normalize(qrot(b.q, j.axis_b));
    }
    // solve_angular_limit_row (solver.h:406-416)
    static void limit_row(BodyT& a, BodyT& b, const M3<T>& Ia, const M3<T>& Ib, const V3<T>& dir, T c, T inv_h, T& acc) {
        T k = dot(dir, mul(Ia, dir)) + dot(dir, mul(Ib, dir));
        if (k <= k_epsilon<T>()) return;
        T cdot = dot(b.w - a.w, dir);
        T lambda = -(cdot + gmax(c, T(0)) * inv_h) / k;
        T nw = gmax(acc + lambda, T(0));
        T delta = nw - acc;
        acc = nw;
        angular_pair(a, b, dir * delta);
    }
    // apply_angular_position_correction (solver.h:456-474)
    static void angular_correction(BodyT& a, BodyT& b, bool ad, bool bd, const M3<T>& Ia, const M3<T>& Ib, const V3<T>& n,
                                   T angle) {
        T k = dot(n, mul(Ia, n)) + dot(n, mul(Ib, n));
        if (k <= k_epsilon<T>()) return;
        V3<T> corr = n * (angle / k);
        if (ad) a.q = rotate_by(a.q, -mul(Ia, corr));
        if (bd) b.q = rotate_by(b.q, mul(Ib, corr));
    }

    // wake_jointed_islands (solver.h:887-901)
    void wake_jointed() {
        bool changed = true;
        for (size_t pass = 0; changed && pass <= joints.size(); ++pass) {
            changed = false;
            for (auto& j : joints) {
                if (!j.valid || !j.enabled) continue;
                BodyT& a = bodies[j.a];
                BodyT& b = bodies[j.b];
                if (a.type != 2 || b.type != 2 || a.awake == b.awake) continue;
                BodyT& s = a.awake ? b : a;
                s.awake = true;
                s.sleep_timer = T(0);
                changed = true;
            }
        }
    }

    // warm_start_joints (solver.h:483-516)
    void warm_start() {
        for (auto& j : joints) {
            if (!j.valid || !j.enabled) continue;
            j.limit_impulse = 0;
            j.swing_impulse = 0;
            BodyT& a = bodies[j.a];
            BodyT& b = bodies[j.b];
            bool aa = act(a), ba = act(b);
            if (!aa && !ba) continue;
            V3<T> ra = qrot(a.q, j.anchor_a), rb = qrot(b.q, j.anchor_b);
            V3<T> pi = j.point_impulse;
            if (dot(pi, pi) > T(0)) {
                if (aa) { a.v -= pi * a.inv_mass; a.w -= mul(a.iiw, cross(ra, pi)); }
                if (ba) { b.v += pi * b.inv_mass; b.w += mul(b.iiw, cross(rb, pi)); }
            }
            if (j.type != J_HINGE) continue;
            V3<T> wa = normalize(qrot(a.q, j.axis_a));
            V3<T> p, q;
            basis(wa, p, q);
            V3<T> ai = p * j.axis_impulse[0] + q * j.axis_impulse[1];
            if (dot(ai, ai) > T(0)) angular_pair(a, b, ai);
        }
    }

    static void basis(const V3<T>& n, V3<T>& t1, V3<T>& t2) {  // util/math.h:62-68
        T sign = n.z >= T(0) ? T(1) : T(-1);
        T a = T(-1) / (sign + n.z);
        T b = n.x * n.y * a;
        t1 = V3<T>(T(1) + sign * n.x * n.x * a, sign * b, -sign * n.x);
        t2 = V3<T>(b, sign + n.y * n.y * a, -n.y);
    }

    // one joint of solve_joint_velocity_pass (solver.h:548-632)
    void velocity_joint(Joint<T>& j, T inv_h) {
        if (!j.valid || !j.enabled) return;
        BodyT& a = bodies[j.a];
        BodyT& b = bodies[j.b];
        bool ad = act(a), bd = act(b);
        if (!ad && !bd) return;
        T ima = ad ? a.inv_mass : T(0), imb = bd ? b.inv_mass : T(0);
        M3<T> Ia = ad ? a.iiw : M3<T>(T(0));
        M3<T> Ib = bd ? b.iiw : M3<T>(T(0));
        if (j.type == J_HINGE) {
            V3<T> wa = normalize(qrot(a.q, j.axis_a));
            V3<T> p, q;
            basis(wa, p, q);
            const V3<T> perp[2] = {p, q};
            for (int k = 0; k < 2; ++k) {
                const V3<T>& n = perp[k];
                T keff = dot(n, mul(Ia, n)) + dot(n, mul(Ib, n));
                if (keff <= k_epsilon<T>()) continue;
                T cdot = dot(b.w - a.w, n);
                T lambda = -cdot / keff;
                j.axis_impulse[k] += lambda;
                angular_pair(a, b, n * lambda);
            }
            if (j.use_limits) {
                T angle = hinge_angle(a.q, b.q, j.axis_a, j.rest, mode);
                T sign, c;
                nearest_limit_side(angle, j.min_angle, j.max_angle, sign, c);
                limit_row(a, b, Ia, Ib, wa * sign, c, inv_h, j.limit_impulse);
            }
        } else if (j.type == J_CONE_TWIST) {
            V3<T> ta, tb;
            cone_axes(a, b, j, ta, tb);
            if (j.has_swing_limit()) {
                V3<T> sw;
                T swing = cone_swing(ta, tb, sw, mode);
                limit_row(a, b, Ia, Ib, -sw, j.swing_limit - swing, inv_h, j.swing_impulse);
            }
            if (j.has_twist_limit()) {
                T angle = hinge_angle(a.q, b.q, j.axis_a, j.rest, mode);
                V3<T> ax = ta + tb;
                T len = length(ax);
                ax = len > T(1e-6f) ? ax / len : ta;
                T sign, c;
                nearest_limit_side(angle, j.twist_min, j.twist_max, sign, c);
                limit_row(a, b, Ia, Ib, ax * sign, c, inv_h, j.limit_impulse);
            }
        }
        V3<T> ra = qrot(a.q, j.anchor_a), rb = qrot(b.q, j.anchor_b);
        M3<T> K;
        if (point_block_mass(ra, rb, ima, imb, Ia, Ib, K)) {
            V3<T> va = a.v + cross(a.w, ra), vb = b.v + cross(b.w, rb);
            V3<T> imp = mul(m3_inverse(K), va - vb);
            j.point_impulse += imp;
            if (ad) { a.v -= imp * ima; a.w -= mul(Ia, cross(ra, imp)); }
            if (bd) { b.v += imp * imb; b.w += mul(Ib, cross(rb, imp)); }
        }
    }
    void velocity_pass(T h, bool reverse) {
        const T inv_h = h > T(0) ? T(1) / h : T(0);
        const size_t n = joints.size();
        for (size_t i = 0; i < n; ++i) velocity_joint(joints[reverse ? n - 1 - i : i], inv_h);
    }

    // one joint of solve_joint_position (solver.h:652-747)
    template <class Cfg> void position_joint(Joint<T>& j, const Cfg& cfg) {
        if (!j.valid || !j.enabled) return;
        BodyT& a = bodies[j.a];
        BodyT& b = bodies[j.b];
        bool ad = act(a), bd = act(b);
        if (!ad && !bd) return;
        T ima = ad ? a.inv_mass : T(0), imb = bd ? b.inv_mass : T(0);
        M3<T> Ia = ad ? a.iiw : M3<T>(T(0));
        M3<T> Ib = bd ? b.iiw : M3<T>(T(0));
        if (j.type == J_HINGE) {
            V3<T> wa = normalize(qrot(a.q, j.axis_a));
            V3<T> wb = normalize(qrot(b.q, j.axis_b));
            V3<T> e = cross(wa, wb);
            T el = length(e);
            if (el > cfg.slop) {
                V3<T> n = e / el;
                T mag = gmin((el - cfg.slop) * cfg.pos_corr, cfg.max_corr);
                angular_correction(a, b, ad, bd, Ia, Ib, n, -mag);
            }
            if (j.use_limits) {
                T angle = hinge_angle(a.q, b.q, j.axis_a, j.rest, mode);
                T over = angle < j.min_angle ? angle - j.min_angle : (angle > j.max_angle ? angle - j.max_angle : T(0));
                if (std::abs(over) > cfg.slop) {
                    T mag = gmin((std::abs(over) - cfg.slop) * cfg.pos_corr, cfg.max_corr);
                    V3<T> ax = normalize(qrot(a.q, j.axis_a));
                    angular_correction(a, b, ad, bd, Ia, Ib, ax, over > T(0) ? -mag : mag);
                }
            }
        } else if (j.type == J_CONE_TWIST) {
            V3<T> ta, tb;
            cone_axes(a, b, j, ta, tb);
            if (j.has_swing_limit()) {
                V3<T> sw;
                T over = cone_swing(ta, tb, sw, mode) - j.swing_limit;
                if (over > cfg.slop) {
                    T mag = gmin((over - cfg.slop) * cfg.pos_corr, cfg.max_corr);
                    angular_correction(a, b, ad, bd, Ia, Ib, sw, -mag);
                    cone_axes(a, b, j, ta, tb);
                }
            }
            if (j.has_twist_limit()) {
                T angle = hinge_angle(a.q, b.q, j.axis_a, j.rest, mode);
                T over = angle < j.twist_min ? angle - j.twist_min : (angle > j.twist_max ? angle - j.twist_max : T(0));
                if (std::abs(over) > cfg.slop) {
                    T mag = gmin((std::abs(over) - cfg.slop) * cfg.pos_corr, cfg.max_corr);
                    V3<T> ax = ta + tb;
                    T len = length(ax);
                    ax = len > T(1e-6f) ? ax / len : ta;
                    angular_correction(a, b, ad, bd, Ia, Ib, ax, over > T(0) ? -mag : mag);
                }
            }
        }
        V3<T> ra = qrot(a.q, j.anchor_a), rb = qrot(b.q, j.anchor_b);
        V3<T> err = (b.pos + rb) - (a.pos + ra);
        T el = length(err);
        M3<T> K;
        if (el > cfg.slop && point_block_mass(ra, rb, ima, imb, Ia, Ib, K)) {
            T mag = gmin((el - cfg.slop) * cfg.pos_corr, cfg.max_corr);
            V3<T> p = mul(m3_inverse(K), (err / el) * mag);
            if (ad) {
                a.pos += p * ima;
                a.q = rotate_by(a.q, mul(Ia, cross(ra, p)));
            }
            if (bd) {
                b.pos -= p * imb;
                b.q = rotate_by(b.q, -mul(Ib, cross(rb, p)));
            }
        }
    }
    template <class Cfg> void position_pass(const Cfg& cfg) {
        for (auto& j : joints) position_joint(j, cfg);
    }
};

}  // namespace pxo
