// Convex-hull narrowphase on the GPU (collision/sat_hull.h): hull against hull, box, sphere and
// capsule, with the reference's SAT axes, tie rules, clipping and feature ids. One thread per
// pair, as for the other shape types; a hull's world vertices are recomputed where the reference
// reads its per-call vertex array (the same arithmetic, so the same values).
//
// The hull table (uploaded by pxg_set_hulls): per hull its vertex, face and edge ranges and local
// bounds; per face its first vertex-list entry and vertex count. Faces have at most 8 vertices,
// so a clipped polygon (at most the incident face's vertices plus one per side plane) has at most
// 16, and keeping the 4 deepest of at most 16 candidates is libstdc++'s std::sort on 16 elements
// or fewer: an insertion sort, which keeps equal depths in clip order.
#pragma once
#include "pxg_device.h"

namespace pxd {

constexpr int SH_HULL = 3;  // pxg_shape_type PXG_CONVEX_HULL
constexpr int kHullFaceMax = 8;   // vertices per face
constexpr int kHullClip = 16;     // clipped polygon vertices

struct HullTab {
    const uint4* meta;    // per hull: vertex offset, vertex count, face offset, face count
    const uint4* meta2;   // per hull: edge offset, edge count
    const float4* lb;     // per hull: local bounds min, max
    const float4* cen;    // per hull: centroid, volume
    const float4* v;      // vertices
    const float4* pl;     // face planes: outward unit normal, offset
    const uint2* face;    // per face: first entry in fv, vertex count
    const uint32_t* fv;   // face vertex lists (hull-local vertex indices)
    const uint2* e;       // edges (hull-local vertex indices)
};

struct PHull {  // sat_hull.h:47-60: a hull posed in the world
    uint32_t v0, nv, f0, nf, e0, ne;
    v3 pos;
    quat rot;
    float scale;
};
PXD PHull posed_hull(const HullTab& t, uint32_t k, v3 pos, quat rot, const ShapeGeom& s) {  // sat_hull.h:65-72
    PHull h;
    uint4 m = t.meta[k], m2 = t.meta2[k];
    h.v0 = m.x; h.nv = m.y; h.f0 = m.z; h.nf = m.w; h.e0 = m2.x; h.ne = m2.y;
    h.pos = pos + qrot(rot, s.lc);
    h.rot = qmul(rot, s.lr);
    h.scale = s.radius;
    return h;
}
PXD v3 hvtx(const HullTab& t, const PHull& h, uint32_t i) { return h.pos + qrot(h.rot, ld3(t.v[h.v0 + i]) * h.scale); }
PXD v3 hfn(const HullTab& t, const PHull& h, uint32_t f) { return qrot(h.rot, ld3(t.pl[h.f0 + f])); }
// the face's first vertex (the plane the reference measures depths from)
PXD v3 hface_v0(const HullTab& t, const PHull& h, uint32_t f) { return hvtx(t, h, t.fv[t.face[h.f0 + f].x]); }

// detail::project_ (sat_hull.h:84-92) of a hull's world vertices and of a box's 8 corners
PXD void hproject(const HullTab& t, const PHull& h, v3 axis, float& lo, float& hi) {
    lo = 3.402823466e+38f;
    hi = -3.402823466e+38f;
    for (uint32_t i = 0; i < h.nv; ++i) {
        float d = dot(hvtx(t, h, i), axis);
        lo = gmin(lo, d);
        hi = gmax(hi, d);
    }
}
PXD void bproject(const v3 c[8], v3 axis, float& lo, float& hi) {
    lo = 3.402823466e+38f;
    hi = -3.402823466e+38f;
    for (int i = 0; i < 8; ++i) {
        float d = dot(c[i], axis);
        lo = gmin(lo, d);
        hi = gmax(hi, d);
    }
}

// detail::clip_polygon_n_ (sat_hull.h:129-145)
PXD void hclip(const v3* in, int nin, v3 pn, float pd, v3* out, int& nout) {
    nout = 0;
    if (nin == 0) return;
    for (int i = 0; i < nin; ++i) {
        v3 cur = in[i];
        v3 prev = in[(i - 1 + nin) % nin];
        float dc = dot(pn, cur) - pd;
        float dp = dot(pn, prev) - pd;
        bool ci = dc <= 0.0f, pi = dp <= 0.0f;
        if (ci != pi) {
            float tt = dp / (dp - dc);
            if (nout < kHullClip) out[nout++] = prev + tt * (cur - prev);
        }
        if (ci && nout < kHullClip) out[nout++] = cur;
    }
}
PXD v3 outward_edge_normal(v3 edge, v3 fn) {  // sat_hull.h:157-161
    v3 n = cross(edge, fn);
    float l = len(n);
    return l > 1e-12f ? n / l : fn;
}
// clip `poly` (count vertices) by the side planes of hull face f (sat_hull.h:165-178)
PXD void hclip_by_face(const HullTab& t, const PHull& h, uint32_t f, v3* poly, v3* tmp, int& count) {
    const uint2 fc = t.face[h.f0 + f];
    const v3 normal = hfn(t, h, f);
    const int n = int(fc.y);
    for (int i = 0; i < n && count > 0; ++i) {
        v3 a = hvtx(t, h, t.fv[fc.x + uint32_t(i)]);
        v3 b = hvtx(t, h, t.fv[fc.x + uint32_t((i + 1) % n)]);
        v3 out = outward_edge_normal(b - a, normal);
        int nc = 0;
        hclip(poly, count, out, dot(out, a), tmp, nc);
        for (int k = 0; k < nc; ++k) poly[k] = tmp[k];
        count = nc;
    }
}
PXD int hface_poly(const HullTab& t, const PHull& h, uint32_t f, v3* out) {  // sat_hull.h:181-187
    const uint2 fc = t.face[h.f0 + f];
    for (uint32_t i = 0; i < fc.y; ++i) out[i] = hvtx(t, h, t.fv[fc.x + i]);
    return int(fc.y);
}
// the kept contact points of a clipped polygon below the reference plane (sat_hull.h:355-375):
// deeper than -epsilon, at most 4, deepest first, equal depths in clip order
PXD int hkeep(const v3* poly, int count, v3 rn, float ref_d, v3* pos, float* pen) {
    int kept = 0;
    for (int i = 0; i < count; ++i) {
        float p = -(dot(rn, poly[i]) - ref_d);
        if (p > -kEps) { pos[kept] = poly[i]; pen[kept] = gmax(p, 0.0f); ++kept; }
    }
    if (kept > 4) {
        for (int i = 1; i < kept; ++i) {
            v3 cp = pos[i];
            float cv = pen[i];
            int j = i;
            while (j > 0 && cv > pen[j - 1]) { pos[j] = pos[j - 1]; pen[j] = pen[j - 1]; --j; }
            pos[j] = cp;
            pen[j] = cv;
        }
        kept = 4;
    }
    return kept;
}

// sat_test_hull_hull + generate_hull_hull_contacts (sat_hull.h:223-380); normal from a to b
__device__ __noinline__ bool hull_hull(const HullTab& t, const PHull& a, const PHull& b, Man& out) {
    float best = 3.402823466e+38f;
    bool found = false;
    int kind = 0, fa = -1, fb = -1, ea = -1, eb = -1;
    float rpen = 0.0f;
    v3 rn = mk(0.0f, 0.0f, 1.0f);
    for (uint32_t f = 0; f < a.nf; ++f) {  // A's faces, one-sided (sat_hull.h:117-123)
        v3 nr = hfn(t, a, f);
        float l2 = dot(nr, nr);
        if (l2 < 1e-10f) continue;
        v3 n = nr / sqrtf(l2);
        float alo, ahi, blo, bhi;
        hproject(t, a, n, alo, ahi);
        hproject(t, b, n, blo, bhi);
        float ov = ahi - blo;
        if (!(ov >= 0.0f)) return false;
        if (ov < best) { best = ov; rpen = ov; rn = n; kind = 0; fa = int(f); fb = -1; ea = -1; eb = -1; found = true; }
    }
    for (uint32_t f = 0; f < b.nf; ++f) {  // B's faces
        v3 nr = hfn(t, b, f);
        float l2 = dot(nr, nr);
        if (l2 < 1e-10f) continue;
        v3 n = nr / sqrtf(l2);
        float alo, ahi, blo, bhi;
        hproject(t, b, n, blo, bhi);
        hproject(t, a, n, alo, ahi);
        float ov = bhi - alo;
        if (!(ov >= 0.0f)) return false;
        if (ov < best) { best = ov; rpen = ov; rn = -n; kind = 1; fa = -1; fb = int(f); ea = -1; eb = -1; found = true; }
    }
    for (uint32_t i = 0; i < a.ne; ++i) {
        uint2 e = t.e[a.e0 + i];
        v3 ed = hvtx(t, a, e.y) - hvtx(t, a, e.x);
        for (uint32_t j = 0; j < b.ne; ++j) {
            uint2 g = t.e[b.e0 + j];
            v3 gd = hvtx(t, b, g.y) - hvtx(t, b, g.x);
            v3 ar = cross(ed, gd);
            float l2 = dot(ar, ar);
            if (l2 < 1e-10f) continue;
            v3 axis = ar / sqrtf(l2);
            float alo, ahi, blo, bhi;
            hproject(t, a, axis, alo, ahi);
            hproject(t, b, axis, blo, bhi);
            float ov = gmin(ahi - blo, bhi - alo);
            if (ov < 0.0f) return false;
            float score = ov + 0.01f;
            if (score < best) {
                best = score; rpen = ov;
                rn = (alo + ahi) >= (blo + bhi) ? -axis : axis;
                kind = 2; fa = -1; fb = -1; ea = int(i); eb = int(j);
                found = true;
            }
        }
    }
    if (!found) return false;
    out.count = 0;
    out.n = rn;
    if (kind == 2) {
        uint2 e = t.e[a.e0 + uint32_t(ea)], g = t.e[b.e0 + uint32_t(eb)];
        v3 c1, c2;
        seg_seg(hvtx(t, a, e.x), hvtx(t, a, e.y), hvtx(t, b, g.x), hvtx(t, b, g.y), c1, c2);
        add_point(out, 0.5f * (c1 + c2), rpen, 0x50000000u | (uint32_t(ea) << 16) | uint32_t(eb));
        return true;
    }
    const bool ref_a = kind == 0;
    const PHull& ref = ref_a ? a : b;
    const PHull& inc = ref_a ? b : a;
    const uint32_t rf = uint32_t(ref_a ? fa : fb);
    const v3 refn = hfn(t, ref, rf);
    uint32_t inf = 0;
    float most = 3.402823466e+38f;
    for (uint32_t f = 0; f < inc.nf; ++f) {
        float dp = dot(hfn(t, inc, f), refn);
        if (dp < most) { most = dp; inf = f; }
    }
    v3 poly[kHullClip], tmp[kHullClip];
    int count = hface_poly(t, inc, inf, poly);
    hclip_by_face(t, ref, rf, poly, tmp, count);
    if (count == 0) return false;
    const float ref_d = dot(refn, hface_v0(t, ref, rf));
    v3 pos[kHullClip];
    float pen[kHullClip];
    const int kept = hkeep(poly, count, refn, ref_d, pos, pen);
    for (int i = 0; i < kept; ++i)
        add_point(out, pos[i], pen[i], 0x40000000u | (rf << 16) | (inf << 8) | uint32_t(i & 0xFF));
    return out.count > 0;
}

// generate_hull_box_contacts (sat_hull.h:390-578); normal from the hull to the box
__device__ __noinline__ bool hull_box(const HullTab& t, const PHull& a, const OBB& b, Man& out) {
    v3 bax[3];
    axes(b, bax);
    v3 bc[8];
    int k = 0;
    for (int sx = -1; sx <= 1; sx += 2)
        for (int sy = -1; sy <= 1; sy += 2)
            for (int sz = -1; sz <= 1; sz += 2)
                bc[k++] = b.c + (float(sx) * b.he.x) * bax[0] + (float(sy) * b.he.y) * bax[1] + (float(sz) * b.he.z) * bax[2];
    float best = 3.402823466e+38f;
    bool found = false;
    int kind = 0, fa = -1, ba = -1, ea = -1, eb = -1;
    float bpen = 0.0f;
    v3 bn = mk(0.0f, 0.0f, 1.0f);
    for (uint32_t f = 0; f < a.nf; ++f) {  // the hull's faces, one-sided
        v3 n = hfn(t, a, f);
        float alo, ahi, blo, bhi;
        hproject(t, a, n, alo, ahi);
        bproject(bc, n, blo, bhi);
        float ov = ahi - blo;
        if (!(ov >= 0.0f)) return false;
        if (ov < best) { best = ov; bpen = ov; bn = n; kind = 0; fa = int(f); ba = -1; ea = -1; eb = -1; found = true; }
    }
    // two-sided axes: the box's 3, then the hull's edges crossed with them (edge bias 0.01)
    auto consider = [&](v3 ar, float bias, int k2, int b_a, int e_a, int e_b) -> bool {
        float l2 = dot(ar, ar);
        if (l2 < 1e-10f) return true;
        v3 axis = ar / sqrtf(l2);
        float alo, ahi, blo, bhi;
        hproject(t, a, axis, alo, ahi);
        bproject(bc, axis, blo, bhi);
        float ov = gmin(ahi - blo, bhi - alo);
        if (ov < 0.0f) return false;
        float score = ov + bias;
        if (score < best) {
            best = score; bpen = ov;
            bn = (alo + ahi) >= (blo + bhi) ? -axis : axis;
            kind = k2; fa = -1; ba = b_a; ea = e_a; eb = e_b;
            found = true;
        }
        return true;
    };
    for (int i = 0; i < 3; ++i)
        if (!consider(bax[i], 0.0f, 1, i, -1, -1)) return false;
    for (uint32_t i = 0; i < a.ne; ++i) {
        uint2 e = t.e[a.e0 + i];
        v3 ed = hvtx(t, a, e.y) - hvtx(t, a, e.x);
        for (int j = 0; j < 3; ++j)
            if (!consider(cross(ed, bax[j]), 0.01f, 2, -1, int(i), j)) return false;
    }
    if (!found) return false;
    out.count = 0;
    out.n = bn;
    if (kind == 2) {
        uint2 e = t.e[a.e0 + uint32_t(ea)];
        v3 e0 = hvtx(t, a, e.x), e1 = hvtx(t, a, e.y);
        v3 mid = 0.5f * (e0 + e1);
        v3 d = mid - b.c;
        int aj = (eb + 1) % 3, ak = (eb + 2) % 3;
        float sj = dot(d, bax[aj]) >= 0.0f ? 1.0f : -1.0f;
        float sk = dot(d, bax[ak]) >= 0.0f ? 1.0f : -1.0f;
        v3 ec = b.c + bax[aj] * (sj * comp(b.he, aj)) + bax[ak] * (sk * comp(b.he, ak));
        v3 g0 = ec - bax[eb] * comp(b.he, eb), g1 = ec + bax[eb] * comp(b.he, eb);
        v3 c1, c2;
        seg_seg(e0, e1, g0, g1, c1, c2);
        add_point(out, 0.5f * (c1 + c2), bpen, 0x60000000u | (uint32_t(ea) << 8) | uint32_t(eb));
        return true;
    }
    v3 poly[kHullClip], tmp[kHullClip], pos[kHullClip];
    float pen[kHullClip];
    if (kind == 0) {  // the hull's face is the reference, the box's most anti-parallel face incident
        const v3 rn = hfn(t, a, uint32_t(fa));
        int inc_axis = 0;
        float inc_sign = 1.0f, most = 3.402823466e+38f;
        for (int ax = 0; ax < 3; ++ax)
            for (int si = 0; si < 2; ++si) {
                float s = si == 0 ? 1.0f : -1.0f;
                float dp = dot(bax[ax] * s, rn);
                if (dp < most) { most = dp; inc_axis = ax; inc_sign = s; }
            }
        v3 inn;
        obb_face(b, bax, inc_axis, inc_sign, poly, inn);
        int count = 4;
        hclip_by_face(t, a, uint32_t(fa), poly, tmp, count);
        if (count == 0) return false;
        const float ref_d = dot(rn, hface_v0(t, a, uint32_t(fa)));
        const int kept = hkeep(poly, count, rn, ref_d, pos, pen);
        for (int i = 0; i < kept; ++i) add_point(out, pos[i], pen[i], 0x40000000u | (uint32_t(fa) << 16) | uint32_t(i & 0xFF));
        return out.count > 0;
    }
    // the box's face is the reference, the hull's most anti-parallel face incident
    const v3 desired = -bn;
    const float ref_sign = dot(bax[ba], desired) >= 0.0f ? 1.0f : -1.0f;
    v3 rc[4], rn;
    obb_face(b, bax, ba, ref_sign, rc, rn);
    uint32_t inf = 0;
    float most = 3.402823466e+38f;
    for (uint32_t f = 0; f < a.nf; ++f) {
        float dp = dot(hfn(t, a, f), rn);
        if (dp < most) { most = dp; inf = f; }
    }
    int count = hface_poly(t, a, inf, poly);
    const v3 u = bax[(ba + 1) % 3], v = bax[(ba + 2) % 3];
    const float hu = comp(b.he, (ba + 1) % 3), hv = comp(b.he, (ba + 2) % 3);
    const v3 pn[4] = {u, -u, v, -v};
    const float pd[4] = {dot(u, b.c) + hu, dot(-u, b.c) + hu, dot(v, b.c) + hv, dot(-v, b.c) + hv};
    for (int p = 0; p < 4 && count > 0; ++p) {  // obb_side_planes (clip.h:102-118)
        int nc = 0;
        hclip(poly, count, pn[p], pd[p], tmp, nc);
        for (int i = 0; i < nc; ++i) poly[i] = tmp[i];
        count = nc;
    }
    if (count == 0) return false;
    const float ref_d = dot(rn, rc[0]);
    const int kept = hkeep(poly, count, rn, ref_d, pos, pen);
    for (int i = 0; i < kept; ++i)
        add_point(out, pos[i], pen[i], 0x70000000u | (uint32_t(ba) << 16) | (inf << 8) | uint32_t(i & 0xFF));
    return out.count > 0;
}

// closest_point_on_convex_polygon_ (sat_hull.h:594-625)
PXD v3 closest_on_polygon(const v3* poly, int n, v3 normal, v3 point) {
    const float pd = dot(normal, poly[0]);
    const v3 proj = point - normal * (dot(normal, point) - pd);
    bool inside = true;
    for (int i = 0; i < n && inside; ++i) {
        v3 a = poly[i], b = poly[(i + 1) % n];
        v3 out = outward_edge_normal(b - a, normal);
        if (dot(out, proj - a) > 1e-6f) inside = false;
    }
    if (inside) return proj;
    v3 best = poly[0];
    float bd = 3.402823466e+38f;
    for (int i = 0; i < n; ++i) {
        v3 a = poly[i], b = poly[(i + 1) % n];
        v3 ab = b - a;
        float l2 = dot(ab, ab);
        float tt = l2 > 1e-12f ? sclamp(dot(proj - a, ab) / l2, 0.0f, 1.0f) : 0.0f;
        v3 p = a + ab * tt;
        float d2 = dot(proj - p, proj - p);
        if (d2 < bd) { bd = d2; best = p; }
    }
    return best;
}
// hull_vs_sphere (sat_hull.h:647-687): normal from the hull toward the sphere (outward through the
// nearest face when the centre is inside)
__device__ __noinline__ bool hull_sphere(const HullTab& t, const PHull& a, const Sph& b, v3& normal, float& pen, v3& point) {
    uint32_t bf = 0;
    float bd = -3.402823466e+38f;
    bool any_pos = false;
    for (uint32_t f = 0; f < a.nf; ++f) {
        v3 n = hfn(t, a, f);
        float d = dot(n, hface_v0(t, a, f));
        float dist = dot(n, b.c) - d;
        if (dist > 0.0f) any_pos = true;
        if (dist > bd) { bd = dist; bf = f; }
    }
    const v3 n = hfn(t, a, bf);
    v3 poly[kHullFaceMax];
    const int count = hface_poly(t, a, bf, poly);
    if (!any_pos) {
        point = closest_on_polygon(poly, count, n, b.c);
        v3 diff = point - b.c;
        float dist = len(diff);
        normal = dist > kEps ? diff / dist : n;
        pen = b.r + (-bd);
        return true;
    }
    point = closest_on_polygon(poly, count, n, b.c);
    v3 diff = b.c - point;
    float dist = len(diff);
    if (dist > b.r) return false;
    normal = dist > kEps ? diff / dist : n;
    pen = b.r - dist;
    return true;
}
// hull_vs_capsule (sat_hull.h:698-722): spheres at both ends and the middle of the core segment
__device__ __noinline__ bool hull_capsule(const HullTab& t, const PHull& a, const Cap& b, Man& out) {
    out.count = 0;
    const v3 samples[3] = {b.a, b.b, 0.5f * (b.a + b.b)};
    v3 bn = mk(0.0f, 0.0f, 1.0f);
    float bp = 0.0f;
    bool hit = false;
    for (int i = 0; i < 3; ++i) {
        Sph probe;
        probe.c = samples[i];
        probe.r = b.r;
        v3 n, p;
        float pen;
        if (hull_sphere(t, a, probe, n, pen, p)) {
            add_point(out, p, pen, 0x80000000u | uint32_t(i));
            if (pen > bp) { bp = pen; bn = n; }
            hit = true;
        }
    }
    if (!hit) return false;
    out.n = bn;
    return true;
}

// narrowphase.h:271-313: the hull branches of generate_contacts (no speculative contacts)
PXD bool hull_contacts(const HullTab& t, const ShapeGeom& sa, v3 pa, quat ra, const ShapeGeom& sb, v3 pb, quat rb, Man& out) {
    out.count = 0;
    out.n = mk(0.0f, 0.0f, 1.0f);
    v3 n, p;
    float pen;
    const int ta = sa.type, tb = sb.type;
    if (ta == SH_HULL && tb == SH_HULL)
        return hull_hull(t, posed_hull(t, sa.hull, pa, ra, sa), posed_hull(t, sb.hull, pb, rb, sb), out);
    if (ta == SH_HULL && tb == SH_BOX) return hull_box(t, posed_hull(t, sa.hull, pa, ra, sa), world_obb(sb, pb, rb), out);
    if (ta == SH_BOX && tb == SH_HULL) {
        if (!hull_box(t, posed_hull(t, sb.hull, pb, rb, sb), world_obb(sa, pa, ra), out)) return false;
        out.n = -out.n;
        return true;
    }
    if (ta == SH_HULL && tb == SH_SPHERE) {
        if (!hull_sphere(t, posed_hull(t, sa.hull, pa, ra, sa), world_sphere(sb, pb, rb), n, pen, p)) return false;
        out.n = n; add_point(out, p, pen, 0); return true;
    }
    if (ta == SH_SPHERE && tb == SH_HULL) {
        if (!hull_sphere(t, posed_hull(t, sb.hull, pb, rb, sb), world_sphere(sa, pa, ra), n, pen, p)) return false;
        out.n = -n; add_point(out, p, pen, 0); return true;
    }
    if (ta == SH_HULL && tb == SH_CAPSULE) return hull_capsule(t, posed_hull(t, sa.hull, pa, ra, sa), world_capsule(sb, pb, rb), out);
    if (ta == SH_CAPSULE && tb == SH_HULL) {
        if (!hull_capsule(t, posed_hull(t, sb.hull, pb, rb, sb), world_capsule(sa, pa, ra), out)) return false;
        out.n = -out.n;
        return true;
    }
    return false;
}

// the world AABB (shape.h:222-225, convex_hull.h:83-87): the local bounds' 8 corners through
// mat4(mat3_cast(rot) * scale) with the posed centre as the translation
PXD void hull_bounds(const HullTab& t, const ShapeGeom& s, v3 pos, quat rot, v3& mn, v3& mx) {
    const v3 c = pos + qrot(rot, s.lc);
    m3 m = mat3_cast(qmul(rot, s.lr));
    m.c0 = m.c0 * s.radius;
    m.c1 = m.c1 * s.radius;
    m.c2 = m.c2 * s.radius;
    const v3 lo = ld3(t.lb[2 * s.hull]), hi = ld3(t.lb[2 * s.hull + 1]);
    mn = splat(3.402823466e+38f);
    mx = splat(-3.402823466e+38f);
    for (int i = 0; i < 8; ++i) {
        float px = (i & 1) ? hi.x : lo.x, py = (i & 2) ? hi.y : lo.y, pz = (i & 4) ? hi.z : lo.z;
        v3 w = mk((m.c0.x * px + m.c1.x * py) + (m.c2.x * pz + c.x * 1.0f),
                  (m.c0.y * px + m.c1.y * py) + (m.c2.y * pz + c.y * 1.0f),
                  (m.c0.z * px + m.c1.z * py) + (m.c2.z * pz + c.z * 1.0f));
        mn = vmin(mn, w);
        mx = vmax(mx, w);
    }
}
PXD float hull_min_extent(const HullTab& t, const ShapeGeom& s) {  // world.h:1417-1420
    const v3 lo = ld3(t.lb[2 * s.hull]), hi = ld3(t.lb[2 * s.hull + 1]);
    const v3 half = (hi - lo) * 0.5f * s.radius;
    return fminf(fminf(half.x, half.y), half.z);
}

}  // namespace pxd
