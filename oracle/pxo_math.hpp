// ORACLE -- test infrastructure only. Nothing in the product may include, link or run this.
//
// Scalar/vector/quaternion/matrix helpers with glm 0.9.9.3's evaluation order for every
// operation the reference's rigid-body step uses (the reference includes <glm/glm.hpp>,
// e.g. physxcoopa/dynamics/solver.h:29, physxcoopa/geometry/obb.h:16-18). glm itself is not in
// this image; the orders below restate glm 0.9.9's published source:
//   dot(vec3)      (x*x' + y*y') + z*z'                     (func_geometric.inl compute_dot)
//   cross          (y*z'-y'*z, z*x'-z'*x, x*y'-x'*y)         (func_geometric.inl cross)
//   normalize(v)   v * (1 / sqrt(dot(v,v)))                  (inversesqrt = 1/sqrt)
//   quat * vec3    v + ((uv*w) + uuv) * 2, uv = cross(q.xyz, v), uuv = cross(q.xyz, uv)
//   quat * quat    Hamilton product, w/x/y/z expressions in glm's term order
//   dot(quat)      (x*x' + y*y') + (z*z' + w*w')             (0.9.9.3 tquat compute_dot)
//   normalize(q)   q * (1/len), identity when len <= 0
//   mat3_cast      glm's qxx..qwz expansion
//   mat3 * vec3    row r: m[0][r]*x + m[1][r]*y + m[2][r]*z
//   mat3 * mat3    R[c][r] = A[0][r]*B[c][0] + A[1][r]*B[c][1] + A[2][r]*B[c][2]
//   mat4 * vec4    (m0*x + m1*y) + (m2*z + m3*w)
// Parity with the real glm binary is unpinned at the bit level (glm cannot be built here); these
// orders are pinned by the closed-form tests in tests/test_oracle_*.py instead.
#pragma once
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <limits>

namespace pxo {

template <class T> struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    explicit V3(T s) : x(s), y(s), z(s) {}
    T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> inline V3<T> operator+(const V3<T>& a, const V3<T>& b) { return V3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline V3<T> operator-(const V3<T>& a, const V3<T>& b) { return V3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline V3<T> operator-(const V3<T>& a) { return V3<T>(-a.x, -a.y, -a.z); }
template <class T> inline V3<T> operator*(const V3<T>& a, T s) { return V3<T>(a.x * s, a.y * s, a.z * s); }
template <class T> inline V3<T> operator*(T s, const V3<T>& a) { return V3<T>(s * a.x, s * a.y, s * a.z); }
template <class T> inline V3<T> operator*(const V3<T>& a, const V3<T>& b) { return V3<T>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> inline V3<T> operator/(const V3<T>& a, T s) { return V3<T>(a.x / s, a.y / s, a.z / s); }
template <class T> inline V3<T>& operator+=(V3<T>& a, const V3<T>& b) { a = a + b; return a; }
template <class T> inline V3<T>& operator-=(V3<T>& a, const V3<T>& b) { a = a - b; return a; }
template <class T> inline V3<T>& operator*=(V3<T>& a, T s) { a = a * s; return a; }

template <class T> inline T dot(const V3<T>& a, const V3<T>& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> inline V3<T> cross(const V3<T>& a, const V3<T>& b) {
    return V3<T>(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
template <class T> inline T length(const V3<T>& v) { return std::sqrt(dot(v, v)); }
template <class T> inline V3<T> normalize(const V3<T>& v) { return v * (T(1) / std::sqrt(dot(v, v))); }
// glm::min/max on scalars: min = (y < x) ? y : x ; max = (x < y) ? y : x
template <class T> inline T gmin(T a, T b) { return (b < a) ? b : a; }
template <class T> inline T gmax(T a, T b) { return (a < b) ? b : a; }
template <class T> inline V3<T> vmin(const V3<T>& a, const V3<T>& b) { return V3<T>(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
template <class T> inline V3<T> vmax(const V3<T>& a, const V3<T>& b) { return V3<T>(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)); }
// std::clamp(v, lo, hi): (v < lo) ? lo : (hi < v) ? hi : v
template <class T> inline T sclamp(T v, T lo, T hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
// glm::clamp(x, lo, hi) = min(max(x, lo), hi)
template <class T> inline T gclamp(T v, T lo, T hi) { return gmin(gmax(v, lo), hi); }
template <class T> inline T gabs(T v) { return v >= T(0) ? v : -v; }

template <class T> struct Q {  // glm storage order x, y, z, w
    T x, y, z, w;
    Q() : x(0), y(0), z(0), w(1) {}
    static Q wxyz(T w_, T x_, T y_, T z_) { Q q; q.x = x_; q.y = y_; q.z = z_; q.w = w_; return q; }
};
template <class T> inline Q<T> qmul(const Q<T>& p, const Q<T>& q) {
    return Q<T>::wxyz(p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z,
                      p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y,
                      p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z,
                      p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x);
}
template <class T> inline V3<T> qrot(const Q<T>& q, const V3<T>& v) {
    V3<T> qv(q.x, q.y, q.z);
    V3<T> uv = cross(qv, v);
    V3<T> uuv = cross(qv, uv);
    return v + ((uv * q.w) + uuv) * T(2);
}
template <class T> inline Q<T> qadd(const Q<T>& a, const Q<T>& b) { return Q<T>::wxyz(a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline Q<T> qscale(const Q<T>& a, T s) { return Q<T>::wxyz(a.w * s, a.x * s, a.y * s, a.z * s); }
template <class T> inline T qdot(const Q<T>& a, const Q<T>& b) { return (a.x * b.x + a.y * b.y) + (a.z * b.z + a.w * b.w); }
template <class T> inline Q<T> qnormalize(const Q<T>& q) {
    T len = std::sqrt(qdot(q, q));
    if (len <= T(0)) return Q<T>::wxyz(1, 0, 0, 0);
    T inv = T(1) / len;
    return Q<T>::wxyz(q.w * inv, q.x * inv, q.y * inv, q.z * inv);
}

template <class T> struct M3 {  // column-major: c[col][row]
    V3<T> c[3];
    M3() {}
    explicit M3(T diag) { c[0] = V3<T>(diag, 0, 0); c[1] = V3<T>(0, diag, 0); c[2] = V3<T>(0, 0, diag); }
};
template <class T> inline M3<T> mat3_cast(const Q<T>& q) {
    M3<T> r(T(1));
    T qxx = q.x * q.x, qyy = q.y * q.y, qzz = q.z * q.z;
    T qxz = q.x * q.z, qxy = q.x * q.y, qyz = q.y * q.z;
    T qwx = q.w * q.x, qwy = q.w * q.y, qwz = q.w * q.z;
    r.c[0].x = T(1) - T(2) * (qyy + qzz);
    r.c[0].y = T(2) * (qxy + qwz);
    r.c[0].z = T(2) * (qxz - qwy);
    r.c[1].x = T(2) * (qxy - qwz);
    r.c[1].y = T(1) - T(2) * (qxx + qzz);
    r.c[1].z = T(2) * (qyz + qwx);
    r.c[2].x = T(2) * (qxz + qwy);
    r.c[2].y = T(2) * (qyz - qwx);
    r.c[2].z = T(1) - T(2) * (qxx + qyy);
    return r;
}
template <class T> inline V3<T> mul(const M3<T>& m, const V3<T>& v) {
    return V3<T>(m.c[0].x * v.x + m.c[1].x * v.y + m.c[2].x * v.z,
                 m.c[0].y * v.x + m.c[1].y * v.y + m.c[2].y * v.z,
                 m.c[0].z * v.x + m.c[1].z * v.y + m.c[2].z * v.z);
}
template <class T> inline M3<T> mul(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row)
            r.c[col][row] = a.c[0][row] * b.c[col][0] + a.c[1][row] * b.c[col][1] + a.c[2][row] * b.c[col][2];
    return r;
}
template <class T> inline M3<T> transpose(const M3<T>& m) {
    M3<T> r;
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) r.c[col][row] = m.c[row][col];
    return r;
}
// vec3(mat4(mat3_cast(q)) with column 3 = (center,1)) * vec4(p, 1))
template <class T> inline V3<T> affine(const M3<T>& m, const V3<T>& center, const V3<T>& p) {
    V3<T> out;
    for (int r = 0; r < 3; ++r) out[r] = (m.c[0][r] * p.x + m.c[1][r] * p.y) + (m.c[2][r] * p.z + center[r] * T(1));
    return out;
}

template <class T> constexpr T k_epsilon() { return T(1e-6f); }

}  // namespace pxo
