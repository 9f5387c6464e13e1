// Device-side geometry, narrowphase and solver-row helpers for gfx950.
//
// Every expression keeps the evaluation order of the reference's glm-based code so the GPU step
// reproduces the reference's FP32 results exactly when compiled with -ffp-contract=off (no fused
// multiply-add) and HIP's default correctly-rounded fp32 division and square root:
//   dot(a,b) = (ax*bx + ay*by) + az*bz; cross in glm's term order; normalize(v) = v * (1/sqrt(v.v));
//   quat*vec3 = v + ((uv*w) + uuv)*2; quat normalize = q * (1/len); mat3*vec3 row-wise sums.
// Citations point at the reference function each helper restates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PXD __device__ __forceinline__

namespace pxd {

struct v3 {
    float x, y, z;
};
PXD v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
PXD v3 splat(float s) { return mk(s, s, s); }
PXD v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PXD v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PXD v3 operator-(v3 a) { return mk(-a.x, -a.y, -a.z); }
PXD v3 operator*(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
PXD v3 operator*(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
PXD v3 operator/(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
PXD float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
PXD void setc(v3& a, int i, float v) { if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v; }
PXD float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
PXD v3 cross(v3 a, v3 b) { return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
PXD float len(v3 a) { return sqrtf(dot(a, a)); }
PXD v3 normalize(v3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
// glm::min / glm::max / std::clamp semantics (ternaries, not fminf: identical for non-NaN input)
PXD float gmin(float a, float b) { return (b < a) ? b : a; }
PXD float gmax(float a, float b) { return (a < b) ? b : a; }
PXD float sclamp(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
PXD float gclamp(float v, float lo, float hi) { return gmin(gmax(v, lo), hi); }
PXD float gabs(float v) { return v >= 0.0f ? v : -v; }
PXD v3 vmin(v3 a, v3 b) { return mk(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
PXD v3 vmax(v3 a, v3 b) { return mk(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)); }
PXD v3 ld3(float4 f) { return mk(f.x, f.y, f.z); }
PXD float4 st3(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
// Load of element i of a float4 array that another workgroup of the same launch may have written
// before the last grid barrier: one device-scope (sc1) 16-byte buffer load, which bypasses the
// CU's L1 (stores write through it). With every such load going through here, a barrier among
// workgroups that share one L2 only has to wait for stores to complete; it needs no L2 write-back
// or invalidation. The base must be the same for the whole wave (it goes into the buffer
// resource); word 3 of the resource is the raw 32-bit buffer format of gfx9-family GPUs.
PXD float4 ld_dev(const float4* base, uint32_t i) {
    __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(base), 0, 0x7FFFFFFF, 0x00020000);
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, static_cast<int>(i * 16u), 0, 16));
}

constexpr float kEps = 1e-6f;  // util::k_epsilon, physxcoopa/util/math.h:19

struct quat {
    float x, y, z, w;
};
PXD quat ldq(float4 f) { quat q; q.x = f.x; q.y = f.y; q.z = f.z; q.w = f.w; return q; }
PXD float4 stq(quat q) { return make_float4(q.x, q.y, q.z, q.w); }
PXD quat qmul(quat p, quat q) {
    quat r;
    r.w = p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z;
    r.x = p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y;
    r.y = p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z;
    r.z = p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x;
    return r;
}
PXD v3 qrot(quat q, v3 v) {
    v3 qv = mk(q.x, q.y, q.z);
    v3 uv = cross(qv, v);
    v3 uuv = cross(qv, uv);
    return v + ((uv * q.w) + uuv) * 2.0f;
}
PXD quat qnormalize(quat q) {
    float l = sqrtf((q.x * q.x + q.y * q.y) + (q.z * q.z + q.w * q.w));
    quat r;
    if (l <= 0.0f) { r.x = 0.0f; r.y = 0.0f; r.z = 0.0f; r.w = 1.0f; return r; }
    float inv = 1.0f / l;
    r.w = q.w * inv; r.x = q.x * inv; r.y = q.y * inv; r.z = q.z * inv;
    return r;
}

// Column-major 3x3: c0, c1, c2 columns.
struct m3 {
    v3 c0, c1, c2;
};
PXD m3 mat3_cast(quat q) {  // glm::mat3_cast
    float qxx = q.x * q.x, qyy = q.y * q.y, qzz = q.z * q.z;
    float qxz = q.x * q.z, qxy = q.x * q.y, qyz = q.y * q.z;
    float qwx = q.w * q.x, qwy = q.w * q.y, qwz = q.w * q.z;
    m3 r;
    r.c0 = mk(1.0f - 2.0f * (qyy + qzz), 2.0f * (qxy + qwz), 2.0f * (qxz - qwy));
    r.c1 = mk(2.0f * (qxy - qwz), 1.0f - 2.0f * (qxx + qzz), 2.0f * (qyz + qwx));
    r.c2 = mk(2.0f * (qxz + qwy), 2.0f * (qyz - qwx), 1.0f - 2.0f * (qxx + qyy));
    return r;
}
PXD v3 mul(const m3& m, v3 v) {
    return mk(m.c0.x * v.x + m.c1.x * v.y + m.c2.x * v.z,
              m.c0.y * v.x + m.c1.y * v.y + m.c2.y * v.z,
              m.c0.z * v.x + m.c1.z * v.y + m.c2.z * v.z);
}
PXD v3 col(const m3& m, int c) { return c == 0 ? m.c0 : (c == 1 ? m.c1 : m.c2); }
// R * diag(d) * R^T with glm's mat3*mat3 term order (dynamics/integrator.h:57-68)
PXD m3 world_inertia(quat q, v3 d) {
    m3 r = mat3_cast(q);
    // (R * D)[c][row] = R[0][row]*D[c][0] + R[1][row]*D[c][1] + R[2][row]*D[c][2]
    v3 rd[3];
    for (int c = 0; c < 3; ++c) {
        float d0 = c == 0 ? d.x : 0.0f, d1 = c == 1 ? d.y : 0.0f, d2 = c == 2 ? d.z : 0.0f;
        rd[c] = mk(r.c0.x * d0 + r.c1.x * d1 + r.c2.x * d2,
                   r.c0.y * d0 + r.c1.y * d1 + r.c2.y * d2,
                   r.c0.z * d0 + r.c1.z * d1 + r.c2.z * d2);
    }
    // RT[c][k] = R[k][c]
    v3 rt[3];
    rt[0] = mk(r.c0.x, r.c1.x, r.c2.x);
    rt[1] = mk(r.c0.y, r.c1.y, r.c2.y);
    rt[2] = mk(r.c0.z, r.c1.z, r.c2.z);
    m3 o;
    v3* oc[3] = {&o.c0, &o.c1, &o.c2};
    for (int c = 0; c < 3; ++c) {
        *oc[c] = mk(rd[0].x * rt[c].x + rd[1].x * rt[c].y + rd[2].x * rt[c].z,
                    rd[0].y * rt[c].x + rd[1].y * rt[c].y + rd[2].y * rt[c].z,
                    rd[0].z * rt[c].x + rd[1].z * rt[c].y + rd[2].z * rt[c].z);
    }
    return o;
}

// ---------------------------------------------------------------------------------------------
// shapes and world instancing (collision/shape.h:160-228, geometry/obb.h)
// ---------------------------------------------------------------------------------------------
enum { SH_SPHERE = 0, SH_BOX = 1, SH_CAPSULE = 2 };  // convex hulls (3): pxg_hull.h

struct ShapeGeom {  // one shape's local description, decoded from the packed device record
    int type, caxis;
    v3 lc;
    quat lr;
    float radius, cr, chh;  // a convex hull's uniform scale is in radius
    v3 he;
    uint32_t hull;          // a convex hull's index in the hull table
};
struct OBB { v3 c, he; quat q; };
struct Sph { v3 c; float r; };
struct Cap { v3 a, b; float r; };

PXD void axes(const OBB& o, v3 ax[3]) {
    m3 r = mat3_cast(o.q);
    ax[0] = r.c0; ax[1] = r.c1; ax[2] = r.c2;
}
PXD Sph world_sphere(const ShapeGeom& s, v3 pos, quat rot) { Sph o; o.c = pos + qrot(rot, s.lc); o.r = s.radius; return o; }
PXD OBB world_obb(const ShapeGeom& s, v3 pos, quat rot) { OBB o; o.c = pos + qrot(rot, s.lc); o.he = s.he; o.q = qmul(rot, s.lr); return o; }
PXD Cap world_capsule(const ShapeGeom& s, v3 pos, quat rot) {
    v3 axl = mk(s.caxis == 0 ? 1.0f : 0.0f, s.caxis == 1 ? 1.0f : 0.0f, s.caxis == 2 ? 1.0f : 0.0f);
    v3 axw = qrot(qmul(rot, s.lr), axl);
    v3 c = pos + qrot(rot, s.lc);
    Cap o;
    o.a = c - axw * s.chh;
    o.b = c + axw * s.chh;
    o.r = s.cr;
    return o;
}
// OBB::bounds (geometry/obb.h:39-49): the 8 corners through mat4_cast with the centre as the
// translation column (aabb.h:83-95)
PXD void obb_bounds(const OBB& o, v3& mn, v3& mx) {
    m3 m = mat3_cast(o.q);
    mn = splat(3.402823466e+38f);
    mx = splat(-3.402823466e+38f);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        float px = (i & 1) ? o.he.x : -o.he.x, py = (i & 2) ? o.he.y : -o.he.y, pz = (i & 4) ? o.he.z : -o.he.z;
        v3 w = mk((m.c0.x * px + m.c1.x * py) + (m.c2.x * pz + o.c.x * 1.0f),
                  (m.c0.y * px + m.c1.y * py) + (m.c2.y * pz + o.c.y * 1.0f),
                  (m.c0.z * px + m.c1.z * py) + (m.c2.z * pz + o.c.z * 1.0f));
        mn = vmin(mn, w);
        mx = vmax(mx, w);
    }
}
// Tight world AABB (collision/shape.h:198-228; box via 8 transformed corners, aabb.h:83-95)
PXD void world_bounds(const ShapeGeom& s, v3 pos, quat rot, v3& mn, v3& mx) {
    if (s.type == SH_SPHERE) {
        Sph sp = world_sphere(s, pos, rot);
        mn = sp.c - splat(sp.r);
        mx = sp.c + splat(sp.r);
    } else if (s.type == SH_BOX) {
        obb_bounds(world_obb(s, pos, rot), mn, mx);
    } else {
        Cap c = world_capsule(s, pos, rot);
        mn = vmin(c.a, c.b) - splat(c.r);
        mx = vmax(c.a, c.b) + splat(c.r);
    }
}
PXD float min_extent(const ShapeGeom& s) {  // world.h:1411-1424
    if (s.type == SH_SPHERE) return s.radius;
    if (s.type == SH_BOX) return fminf(fminf(s.he.x, s.he.y), s.he.z);
    return fminf(s.cr, s.chh);
}

PXD v3 closest_point_on_obb(v3 p, const OBB& b) {  // geometry/obb.h:65-75
    v3 ax[3];
    axes(b, ax);
    v3 d = p - b.c;
    float px = sclamp(dot(d, ax[0]), -b.he.x, b.he.x);
    float py = sclamp(dot(d, ax[1]), -b.he.y, b.he.y);
    float pz = sclamp(dot(d, ax[2]), -b.he.z, b.he.z);
    return b.c + ax[0] * px + ax[1] * py + ax[2] * pz;
}
PXD v3 closest_point_on_segment(v3 p, v3 a, v3 b) {  // geometry/capsule.h:30-40
    v3 ab = b - a;
    float l2 = dot(ab, ab);
    if (l2 < kEps) return a;
    float t = dot(p - a, ab) / l2;
    t = sclamp(t, 0.0f, 1.0f);
    return a + ab * t;
}
PXD void seg_seg(v3 p1, v3 q1, v3 p2, v3 q2, v3& c1, v3& c2) {  // geometry/capsule.h:73-134
    v3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    float s, t;
    if (a < kEps && e < kEps) { c1 = p1; c2 = p2; return; }
    if (a < kEps) {
        s = 0.0f;
        t = sclamp(f / e, 0.0f, 1.0f);
    } else {
        float c = dot(d1, r);
        if (e < kEps) {
            t = 0.0f;
            s = sclamp(-c / a, 0.0f, 1.0f);
        } else {
            float b = dot(d1, d2);
            float denom = a * e - b * b;
            s = denom > kEps ? sclamp((b * f - c * e) / denom, 0.0f, 1.0f) : 0.0f;
            t = (b * s + f) / e;
            if (t < 0.0f) { t = 0.0f; s = sclamp(-c / a, 0.0f, 1.0f); }
            else if (t > 1.0f) { t = 1.0f; s = sclamp((b - c) / a, 0.0f, 1.0f); }
        }
    }
    c1 = p1 + d1 * s;
    c2 = p2 + d2 * t;
}

// ---------------------------------------------------------------------------------------------
// contact output
// ---------------------------------------------------------------------------------------------
struct Pt { v3 pos; float pen, sep; uint32_t fid; };
struct Man {
    v3 n;
    Pt p[4];
    int count;
};
PXD void add_point(Man& m, v3 pos, float pen, uint32_t fid) {  // collision/manifold.h:78-84
    if (m.count >= 4) return;
    Pt& q = m.p[m.count++];
    q.pos = pos; q.pen = pen; q.sep = 0.0f; q.fid = fid;
}
PXD void add_spec(Man& m, v3 pos, float sep, uint32_t fid) {  // collision/manifold.h:89-95
    if (m.count >= 4) return;
    Pt& q = m.p[m.count++];
    q.pos = pos; q.pen = 0.0f; q.sep = sep; q.fid = fid;
}

// collision/narrowphase.h:40-52
PXD bool sphere_sphere(const Sph& a, const Sph& b, v3& n, float& pen, v3& pt) {
    v3 d = b.c - a.c;
    float d2 = dot(d, d), rs = a.r + b.r;
    if (d2 > rs * rs) return false;
    float dist = sqrtf(d2);
    n = dist > kEps ? d / dist : mk(0.0f, 0.0f, 1.0f);
    pen = rs - dist;
    pt = a.c + n * a.r;
    return true;
}
// collision/narrowphase.h:112-122
PXD bool sphere_sphere_sep(const Sph& a, const Sph& b, v3& n, float& sep, v3& pt) {
    v3 d = b.c - a.c;
    float dist = len(d), rs = a.r + b.r;
    if (dist <= rs || dist < kEps) return false;
    n = d / dist;
    sep = dist - rs;
    pt = a.c + n * a.r;
    return true;
}
// collision/narrowphase.h:64-101 (inside) and 133-154 (separation, spec=true)
PXD int sphere_box_any(const Sph& s, const OBB& b, bool spec, v3& n, float& depth, v3& pt) {
    v3 ax[3];
    axes(b, ax);
    v3 d = s.c - b.c;
    v3 l = mk(dot(d, ax[0]), dot(d, ax[1]), dot(d, ax[2]));
    bool inside = fabsf(l.x) <= b.he.x && fabsf(l.y) <= b.he.y && fabsf(l.z) <= b.he.z;
    v3 nl, cl;
    if (!inside) {
        cl = mk(gclamp(l.x, -b.he.x, b.he.x), gclamp(l.y, -b.he.y, b.he.y), gclamp(l.z, -b.he.z, b.he.z));
        v3 diff = l - cl;
        float dist = len(diff);
        if (dist > s.r) {
            if (!spec) return 0;
            // sphere_vs_box_separation: dist > r > ... (dist <= r was excluded above)
            depth = dist - s.r;
            nl = diff / dist;
            n = ax[0] * nl.x + ax[1] * nl.y + ax[2] * nl.z;
            pt = b.c + ax[0] * cl.x + ax[1] * cl.y + ax[2] * cl.z;
            return 2;
        }
        depth = s.r - dist;
        nl = dist > kEps ? diff / dist : mk(0.0f, 0.0f, 1.0f);
    } else {
        v3 fd = mk(b.he.x - gabs(l.x), b.he.y - gabs(l.y), b.he.z - gabs(l.z));
        int axis = 0;
        if (fd.y < fd.x) axis = 1;
        if (fd.z < (axis == 0 ? fd.x : fd.y)) axis = 2;
        float sgn = comp(l, axis) >= 0.0f ? 1.0f : -1.0f;
        nl = mk(axis == 0 ? sgn : 0.0f, axis == 1 ? sgn : 0.0f, axis == 2 ? sgn : 0.0f);
        depth = comp(fd, axis) + s.r;
        cl = l;
        setc(cl, axis, sgn * comp(b.he, axis));
    }
    n = ax[0] * nl.x + ax[1] * nl.y + ax[2] * nl.z;
    pt = b.c + ax[0] * cl.x + ax[1] * cl.y + ax[2] * cl.z;
    return 1;
}

// collision/sat.h:75-160
struct Sat { float pen, sep; int axis; v3 n; bool separated; };
PXD bool sat_obb_obb(const OBB& a, const OBB& b, Sat& res, bool allow_sep) {
    v3 aa[3], ba[3];
    axes(a, aa);
    axes(b, ba);
    v3 d = b.c - a.c;
    float best_score = 3.402823466e+38f, best_overlap = 0.0f;
    int best_axis = -1;
    v3 best_n = mk(0.0f, 0.0f, 1.0f);
    bool separated = false;
    float best_sep_overlap = -3.402823466e+38f;
    int best_sep_axis = -1;
    v3 best_sep_n = mk(0.0f, 0.0f, 1.0f);
    for (int id = 0; id < 15; ++id) {
        if (separated && !allow_sep) break;
        v3 raw;
        float bias = 0.0f;
        if (id < 3) raw = aa[id];
        else if (id < 6) raw = ba[id - 3];
        else { raw = cross(aa[(id - 6) / 3], ba[(id - 6) % 3]); bias = 0.01f; }
        float l2 = dot(raw, raw);
        if (l2 < 1e-8f) continue;
        v3 axis = raw / sqrtf(l2);
        float ra = 0.0f, rb = 0.0f;
        ra += a.he.x * fabsf(dot(axis, aa[0]));
        rb += b.he.x * fabsf(dot(axis, ba[0]));
        ra += a.he.y * fabsf(dot(axis, aa[1]));
        rb += b.he.y * fabsf(dot(axis, ba[1]));
        ra += a.he.z * fabsf(dot(axis, aa[2]));
        rb += b.he.z * fabsf(dot(axis, ba[2]));
        float dd = dot(d, axis);
        float overlap = ra + rb - fabsf(dd);
        v3 n = dd < 0.0f ? -axis : axis;
        if (overlap < 0.0f) {
            separated = true;
            if (allow_sep && overlap > best_sep_overlap) { best_sep_overlap = overlap; best_sep_axis = id; best_sep_n = n; }
            continue;
        }
        float score = overlap + bias;
        if (score < best_score) { best_score = score; best_overlap = overlap; best_axis = id; best_n = n; }
    }
    // one exit, every field written by selects on `separated`
    const bool ok = separated ? (allow_sep && best_sep_axis >= 0) : (best_axis >= 0);
    res.separated = separated;
    res.pen = separated ? 0.0f : best_overlap;
    res.sep = separated ? -best_sep_overlap : 0.0f;
    res.axis = separated ? best_sep_axis : best_axis;
    res.n = separated ? best_sep_n : best_n;
    return ok;
}

// Sutherland-Hodgman clip against one plane (collision/clip.h:40-64); at most 9 vertices.
PXD void clip_plane(const v3* in, int nin, v3 pn, float pd, v3* out, int& nout) {
    nout = 0;
    if (nin == 0) return;
    for (int i = 0; i < nin; ++i) {
        v3 cur = in[i];
        v3 prev = in[(i - 1 + nin) % nin];
        float dc = dot(pn, cur) - pd;
        float dp = dot(pn, prev) - pd;
        bool ci = dc <= 0.0f, pi = dp <= 0.0f;
        if (ci != pi) {
            float t = dp / (dp - dc);
            if (nout < 9) out[nout++] = prev + t * (cur - prev);
        }
        if (ci && nout < 9) out[nout++] = cur;
    }
}
PXD void obb_face(const OBB& b, const v3 ax[3], int axis, float sign, v3 c[4], v3& normal) {  // clip.h:75-92
    v3 m = ax[axis], u = ax[(axis + 1) % 3], v = ax[(axis + 2) % 3];
    float hu = comp(b.he, (axis + 1) % 3), hv = comp(b.he, (axis + 2) % 3);
    v3 fc = b.c + m * (sign * comp(b.he, axis));
    c[0] = fc - u * hu - v * hv;
    c[1] = fc + u * hu - v * hv;
    c[2] = fc + u * hu + v * hv;
    c[3] = fc - u * hu + v * hv;
    normal = m * sign;
}

// collision/sat.h:190-331
PXD bool box_box(const OBB& a, const OBB& b, Man& out, bool spec) {
    Sat r;
    if (!sat_obb_obb(a, b, r, spec)) return false;
    out.count = 0;
    out.n = r.n;
    if (r.separated) {
        v3 ca = closest_point_on_obb(b.c, a);
        v3 cb = closest_point_on_obb(a.c, b);
        add_spec(out, 0.5f * (ca + cb), r.sep, 0x2000u | uint32_t(r.axis));
        return true;
    }
    if (r.axis <= 5) {
        bool ref_is_a = r.axis <= 2;
        const OBB& ref = ref_is_a ? a : b;
        const OBB& inc = ref_is_a ? b : a;
        int ref_axis = ref_is_a ? r.axis : r.axis - 3;
        v3 rax[3], iax[3];
        axes(ref, rax);
        axes(inc, iax);
        v3 desired = ref_is_a ? r.n : -r.n;
        float ref_sign = dot(rax[ref_axis], desired) >= 0.0f ? 1.0f : -1.0f;
        v3 rc[4], rn;
        obb_face(ref, rax, ref_axis, ref_sign, rc, rn);
        int inc_axis = 0;
        float inc_sign = 1.0f, most = 3.402823466e+38f;
        for (int k = 0; k < 3; ++k) {
            for (int si = 0; si < 2; ++si) {
                float s = si == 0 ? 1.0f : -1.0f;
                float dp = dot(iax[k] * s, rn);
                if (dp < most) { most = dp; inc_axis = k; inc_sign = s; }
            }
        }
        v3 pa[9], pb[9], inn;
        obb_face(inc, iax, inc_axis, inc_sign, pa, inn);
        int na = 4;
        // side planes of the reference face (clip.h:102-118)
        v3 u = rax[(ref_axis + 1) % 3], v = rax[(ref_axis + 2) % 3];
        float hu = comp(ref.he, (ref_axis + 1) % 3), hv = comp(ref.he, (ref_axis + 2) % 3);
        for (int p = 0; p < 4 && na > 0; ++p) {
            v3 pn = p == 0 ? u : (p == 1 ? -u : (p == 2 ? v : -v));
            float pd = dot(pn, ref.c) + ((p < 2) ? hu : hv);
            int nb = 0;
            clip_plane(pa, na, pn, pd, pb, nb);
            for (int i = 0; i < nb; ++i) pa[i] = pb[i];
            na = nb;
        }
        if (na == 0) return false;
        float ref_d = dot(rn, rc[0]);
        // keep points with penetration > -eps (clamped at 0); stable insertion sort by depth
        // (descending) when more than four survive, as libstdc++ std::sort does for n <= 16
        float cpen[9];
        int kept = 0;
        for (int i = 0; i < na; ++i) {
            float pen = -(dot(rn, pa[i]) - ref_d);
            if (pen > -kEps) { pb[kept] = pa[i]; cpen[kept] = gmax(pen, 0.0f); ++kept; }
        }
        if (kept == 0) return false;
        if (kept > 4) {
            for (int i = 1; i < kept; ++i) {
                v3 cp = pb[i];
                float cd = cpen[i];
                int j = i;
                while (j > 0 && cd > cpen[j - 1]) { pb[j] = pb[j - 1]; cpen[j] = cpen[j - 1]; --j; }
                pb[j] = cp;
                cpen[j] = cd;
            }
            kept = 4;
        }
        uint32_t rf = uint32_t(ref_axis * 2 + (ref_sign < 0.0f ? 1 : 0));
        uint32_t inf = uint32_t(inc_axis * 2 + (inc_sign < 0.0f ? 1 : 0));
        for (int i = 0; i < kept; ++i) add_point(out, pb[i], cpen[i], (rf << 8) | (inf << 4) | uint32_t(i & 0xF));
        return out.count > 0;
    }
    int idx = r.axis - 6, ei = idx / 3, ej = idx % 3;
    v3 aa[3], ba[3];
    axes(a, aa);
    axes(b, ba);
    v3 dab = b.c - a.c;
    int aj = (ei + 1) % 3, ak = (ei + 2) % 3;
    float asj = dot(dab, aa[aj]) >= 0.0f ? 1.0f : -1.0f;
    float ask = dot(dab, aa[ak]) >= 0.0f ? 1.0f : -1.0f;
    v3 eac = a.c + aa[aj] * (asj * comp(a.he, aj)) + aa[ak] * (ask * comp(a.he, ak));
    v3 ea0 = eac - aa[ei] * comp(a.he, ei), ea1 = eac + aa[ei] * comp(a.he, ei);
    v3 dba = -dab;
    int bj = (ej + 1) % 3, bk = (ej + 2) % 3;
    float bsj = dot(dba, ba[bj]) >= 0.0f ? 1.0f : -1.0f;
    float bsk = dot(dba, ba[bk]) >= 0.0f ? 1.0f : -1.0f;
    v3 ebc = b.c + ba[bj] * (bsj * comp(b.he, bj)) + ba[bk] * (bsk * comp(b.he, bk));
    v3 eb0 = ebc - ba[ej] * comp(b.he, ej), eb1 = ebc + ba[ej] * comp(b.he, ej);
    v3 c1, c2;
    seg_seg(ea0, ea1, eb0, eb1, c1, c2);
    add_point(out, 0.5f * (c1 + c2), r.pen, 0x1000u | (uint32_t(ei) << 4) | uint32_t(ej));
    return true;
}

// collision/segment.h:51-63
PXD bool capsule_sphere(const Cap& c, const Sph& s, v3& n, float& pen, v3& pt) {
    v3 cl = closest_point_on_segment(s.c, c.a, c.b);
    v3 d = s.c - cl;
    float d2 = dot(d, d), rs = c.r + s.r;
    if (d2 > rs * rs) return false;
    float dist = sqrtf(d2);
    n = dist > kEps ? d / dist : mk(0.0f, 0.0f, 1.0f);
    pen = rs - dist;
    pt = cl + n * c.r;
    return true;
}
// collision/segment.h:68-81
PXD bool capsule_capsule(const Cap& a, const Cap& b, v3& n, float& pen, v3& pt) {
    v3 c1, c2;
    seg_seg(a.a, a.b, b.a, b.b, c1, c2);
    v3 d = c2 - c1;
    float d2 = dot(d, d), rs = a.r + b.r;
    if (d2 > rs * rs) return false;
    float dist = sqrtf(d2);
    n = dist > kEps ? d / dist : mk(0.0f, 0.0f, 1.0f);
    pen = rs - dist;
    pt = c1 + n * a.r;
    return true;
}
// collision/segment.h:91-136 (closest_points_segment_obb: 8 alternating projections, :39-46)
PXD bool capsule_box(const Cap& cap, const OBB& box, Man& out) {
    v3 os = 0.5f * (cap.a + cap.b), ob;
    for (int i = 0; i < 8; ++i) {
        ob = closest_point_on_obb(os, box);
        os = closest_point_on_segment(ob, cap.a, cap.b);
    }
    v3 delta = os - ob;
    float dist = len(delta);
    if (dist > cap.r) return false;
    v3 pn = dist > kEps ? delta / dist : mk(0.0f, 0.0f, 1.0f);
    float ppen = cap.r - dist;
    out.count = 0;
    out.n = pn;
    v3 ba_ = closest_point_on_obb(cap.a, box);
    v3 bb_ = closest_point_on_obb(cap.b, box);
    float da = len(cap.a - ba_), db = len(cap.b - bb_);
    bool at = da <= cap.r + 0.02f, bt = db <= cap.r + 0.02f;
    if (at && bt) {
        v3 na = cap.a - ba_, nb = cap.b - bb_;
        float nal = len(na), nbl = len(nb);
        v3 n_a = nal > kEps ? na / nal : pn;
        v3 n_b = nbl > kEps ? nb / nbl : pn;
        if (dot(n_a, n_b) > 0.9f) {
            out.n = normalize(n_a + n_b);
            add_point(out, ba_, gmax(cap.r - da, 0.0f), 0);
            add_point(out, bb_, gmax(cap.r - db, 0.0f), 1);
            return true;
        }
    }
    add_point(out, ob, ppen, 0);
    return true;
}

// collision/narrowphase.h:175-335, box/sphere/capsule branches. Returns true on a manifold.
PXD bool generate_contacts(const ShapeGeom& sa, v3 pa, quat ra, const ShapeGeom& sb, v3 pb, quat rb, bool spec,
                           Man& out) {
    out.count = 0;
    out.n = mk(0.0f, 0.0f, 1.0f);
    v3 n, p;
    float depth;
    int ta = sa.type, tb = sb.type;
    if (ta == SH_SPHERE && tb == SH_SPHERE) {
        Sph A = world_sphere(sa, pa, ra), B = world_sphere(sb, pb, rb);
        if (sphere_sphere(A, B, n, depth, p)) { out.n = n; add_point(out, p, depth, 0); return true; }
        if (spec && sphere_sphere_sep(A, B, n, depth, p)) { out.n = n; add_spec(out, p, depth, 0); return true; }
        return false;
    }
    if (ta == SH_SPHERE && tb == SH_BOX) {
        int k = sphere_box_any(world_sphere(sa, pa, ra), world_obb(sb, pb, rb), spec, n, depth, p);
        if (k == 0) return false;
        out.n = -n;
        if (k == 1) add_point(out, p, depth, 0); else add_spec(out, p, depth, 0);
        return true;
    }
    if (ta == SH_BOX && tb == SH_SPHERE) {
        int k = sphere_box_any(world_sphere(sb, pb, rb), world_obb(sa, pa, ra), spec, n, depth, p);
        if (k == 0) return false;
        out.n = n;
        if (k == 1) add_point(out, p, depth, 0); else add_spec(out, p, depth, 0);
        return true;
    }
    if (ta == SH_BOX && tb == SH_BOX) return box_box(world_obb(sa, pa, ra), world_obb(sb, pb, rb), out, spec);
    if (ta == SH_CAPSULE && tb == SH_CAPSULE) {
        if (!capsule_capsule(world_capsule(sa, pa, ra), world_capsule(sb, pb, rb), n, depth, p)) return false;
        out.n = n; add_point(out, p, depth, 0); return true;
    }
    if (ta == SH_CAPSULE && tb == SH_SPHERE) {
        if (!capsule_sphere(world_capsule(sa, pa, ra), world_sphere(sb, pb, rb), n, depth, p)) return false;
        out.n = n; add_point(out, p, depth, 0); return true;
    }
    if (ta == SH_SPHERE && tb == SH_CAPSULE) {
        if (!capsule_sphere(world_capsule(sb, pb, rb), world_sphere(sa, pa, ra), n, depth, p)) return false;
        out.n = -n; add_point(out, p, depth, 0); return true;
    }
    if (ta == SH_CAPSULE && tb == SH_BOX) {
        if (!capsule_box(world_capsule(sa, pa, ra), world_obb(sb, pb, rb), out)) return false;
        out.n = -out.n;
        return true;
    }
    if (ta == SH_BOX && tb == SH_CAPSULE) return capsule_box(world_capsule(sb, pb, rb), world_obb(sa, pa, ra), out);
    return false;
}

// dynamics/physics_material.h:70-79
PXD float combine(float a, float b, int ma, int mb) {
    int m = ma > mb ? ma : mb;
    if (m == 3) return a * b;
    if (m == 2) return gmax(a, b);
    if (m == 1) return gmin(a, b);
    return 0.5f * (a + b);
}

// util/math.h:62-68
PXD void orthonormal_basis(v3 n, v3& t1, v3& t2) {
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    t1 = mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    t2 = mk(b, sign + n.y * n.y * a, -n.y);
}

// dynamics/solver.h:46-51
PXD uint64_t pack_key(uint32_t ia, uint32_t ib, uint32_t fid) {
    uint32_t lo = ia < ib ? ia : ib, hi = ia < ib ? ib : ia;
    return (uint64_t(lo) << 40) ^ (uint64_t(hi) << 12) ^ uint64_t(fid);
}

}  // namespace pxd
