// ORACLE -- test infrastructure only. Nothing in the product may include, link or run this.
//
// Convex hulls: the deterministic builder of geometry/convex_hull.h:188-449 (incremental hull,
// coplanar merge, volume and centroid) and the hull narrowphase of collision/sat_hull.h
// (hull-hull, hull-box, hull-sphere, hull-capsule SAT contacts with their feature ids), restated
// function by function.
#pragma once
#include "pxo_world.hpp"

#include <map>
#include <vector>

namespace pxo {

// geometry::ConvexHull (convex_hull.h:38-94): vertices, merged faces (outward unit normal and plane
// offset), face vertex lists (CCW about the normal), unique edges, volume, centroid, local bounds
template <class T> struct Hull {
    std::vector<V3<T>> v;
    std::vector<V3<T>> pn;
    std::vector<T> pd;
    std::vector<uint32_t> fs;  // face f spans [fs[f], fs[f + 1]) of fv
    std::vector<uint32_t> fv;
    std::vector<std::pair<uint32_t, uint32_t>> e;
    V3<T> centroid{T(0)};
    T volume = 0;
    AABB<T> lb;
    size_t faces() const { return pn.size(); }
    template <class U> Hull<U> as() const {
        Hull<U> o;
        for (auto& x : v) o.v.push_back(V3<U>(U(x.x), U(x.y), U(x.z)));
        for (auto& x : pn) o.pn.push_back(V3<U>(U(x.x), U(x.y), U(x.z)));
        for (auto x : pd) o.pd.push_back(U(x));
        o.fs = fs; o.fv = fv; o.e = e;
        o.centroid = V3<U>(U(centroid.x), U(centroid.y), U(centroid.z));
        o.volume = U(volume);
        o.lb.mn = V3<U>(U(lb.mn.x), U(lb.mn.y), U(lb.mn.z));
        o.lb.mx = V3<U>(U(lb.mx.x), U(lb.mx.y), U(lb.mx.z));
        return o;
    }
};

namespace hull_detail {

struct Face { uint32_t a = 0, b = 0, c = 0; V3<float> n{0.0f}; float d = 0.0f; };

inline void face_plane(Face& f, const std::vector<V3<float>>& p) {  // convex_hull.h:105-110
    V3<float> n = cross(p[f.b] - p[f.a], p[f.c] - p[f.a]);
    float len = length(n);
    f.n = len > 1e-12f ? n / len : V3<float>(0.0f, 0.0f, 1.0f);
    f.d = dot(f.n, p[f.a]);
}

// convex_hull.h:115-146: farthest point from the centroid first, then greedy farthest-from-kept
inline std::vector<V3<float>> farthest_subset(const V3<float>* pts, size_t n, size_t target) {
    std::vector<V3<float>> kept;
    V3<float> c(0.0f);
    for (size_t i = 0; i < n; ++i) c += pts[i];
    c = c / float(n);
    std::vector<bool> used(n, false);
    size_t first = 0;
    float best = -1.0f;
    for (size_t i = 0; i < n; ++i) {
        float d2 = dot(pts[i] - c, pts[i] - c);
        if (d2 > best) { best = d2; first = i; }
    }
    kept.push_back(pts[first]);
    used[first] = true;
    std::vector<float> md(n, std::numeric_limits<float>::max());
    while (kept.size() < target) {
        size_t pick = n;
        float pd = -1.0f;
        for (size_t i = 0; i < n; ++i) {
            if (used[i]) continue;
            V3<float> df = pts[i] - kept.back();
            md[i] = std::min(md[i], dot(df, df));
            if (md[i] > pd) { pd = md[i]; pick = i; }
        }
        if (pick == n) break;
        kept.push_back(pts[pick]);
        used[pick] = true;
    }
    return kept;
}

struct V2 { float x, y; };
// convex_hull.h:151-186: Andrew's monotone chain over the face's own 2-D projection
inline std::vector<uint32_t> hull_2d(const std::vector<V2>& p) {
    std::vector<uint32_t> order(p.size());
    for (uint32_t i = 0; i < p.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        if (p[a].x != p[b].x) return p[a].x < p[b].x;
        return p[a].y < p[b].y;
    });
    order.erase(std::unique(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
                    float dx = p[a].x - p[b].x, dy = p[a].y - p[b].y;
                    return dx * dx + dy * dy < 1e-14f;
                }),
                order.end());
    if (order.size() < 3) return order;
    auto cr = [&](uint32_t o, uint32_t a, uint32_t b) {
        float oax = p[a].x - p[o].x, oay = p[a].y - p[o].y, obx = p[b].x - p[o].x, oby = p[b].y - p[o].y;
        return oax * oby - oay * obx;
    };
    std::vector<uint32_t> lo, up;
    for (uint32_t i : order) {
        while (lo.size() >= 2 && cr(lo[lo.size() - 2], lo.back(), i) <= 0.0f) lo.pop_back();
        lo.push_back(i);
    }
    for (auto it = order.rbegin(); it != order.rend(); ++it) {
        while (up.size() >= 2 && cr(up[up.size() - 2], up.back(), *it) <= 0.0f) up.pop_back();
        up.push_back(*it);
    }
    lo.pop_back();
    up.pop_back();
    lo.insert(lo.end(), up.begin(), up.end());
    return lo;
}

}  // namespace hull_detail

// build_convex_hull (convex_hull.h:188-449), in float as the reference builds it
inline bool build_convex_hull(const V3<float>* points, size_t count, Hull<float>& out, float tol = 1e-5f,
                              size_t max_vertices = 64) {
    using hull_detail::Face;
    out = Hull<float>{};
    if (count < 4) return false;
    std::vector<V3<float>> cloud = count > max_vertices ? hull_detail::farthest_subset(points, count, max_vertices)
                                                       : std::vector<V3<float>>(points, points + count);
    if (cloud.size() < 4) return false;
    size_t p0 = 0;
    for (size_t i = 1; i < cloud.size(); ++i)
        if (cloud[i].x < cloud[p0].x ||
            (cloud[i].x == cloud[p0].x && (cloud[i].y < cloud[p0].y || (cloud[i].y == cloud[p0].y && cloud[i].z < cloud[p0].z))))
            p0 = i;
    size_t p1 = p0;
    float b1 = -1.0f;
    for (size_t i = 0; i < cloud.size(); ++i) {
        float d2 = dot(cloud[i] - cloud[p0], cloud[i] - cloud[p0]);
        if (d2 > b1) { b1 = d2; p1 = i; }
    }
    if (b1 < tol * tol) return false;
    size_t p2 = p0;
    float b2 = -1.0f;
    V3<float> ld = normalize(cloud[p1] - cloud[p0]);
    for (size_t i = 0; i < cloud.size(); ++i) {
        V3<float> t = cloud[i] - cloud[p0];
        V3<float> perp = t - ld * dot(t, ld);
        float d2 = dot(perp, perp);
        if (d2 > b2) { b2 = d2; p2 = i; }
    }
    if (b2 < tol * tol) return false;
    V3<float> sn = normalize(cross(cloud[p1] - cloud[p0], cloud[p2] - cloud[p0]));
    float sd = dot(sn, cloud[p0]);
    size_t p3 = p0;
    float b3 = 0.0f;
    for (size_t i = 0; i < cloud.size(); ++i) {
        float dist = std::abs(dot(sn, cloud[i]) - sd);
        if (dist > b3) { b3 = dist; p3 = i; }
    }
    if (b3 < tol) return false;
    uint32_t ia = uint32_t(p0), ib = uint32_t(p1), ic = uint32_t(p2), id = uint32_t(p3);
    auto make = [&](uint32_t a, uint32_t b, uint32_t c, uint32_t opp) {
        Face f;
        f.a = a; f.b = b; f.c = c;
        hull_detail::face_plane(f, cloud);
        if (dot(f.n, cloud[opp]) - f.d > 0.0f) std::swap(f.b, f.c);
        hull_detail::face_plane(f, cloud);
        return f;
    };
    std::vector<Face> faces{make(ia, ib, ic, id), make(ia, ib, id, ic), make(ib, ic, id, ia), make(ia, ic, id, ib)};
    std::vector<bool> used(cloud.size(), false);
    used[ia] = used[ib] = used[ic] = used[id] = true;
    for (;;) {
        size_t bp = cloud.size();
        float bd = tol;
        for (size_t i = 0; i < cloud.size(); ++i) {
            if (used[i]) continue;
            float far = 0.0f;
            for (const Face& f : faces) far = std::max(far, dot(f.n, cloud[i]) - f.d);
            if (far > bd) { bd = far; bp = i; }
        }
        if (bp == cloud.size()) break;
        uint32_t pt = uint32_t(bp);
        std::vector<bool> vis(faces.size(), false);
        for (size_t i = 0; i < faces.size(); ++i) vis[i] = dot(faces[i].n, cloud[pt]) - faces[i].d > tol;
        std::map<std::pair<uint32_t, uint32_t>, size_t> owner;
        for (size_t i = 0; i < faces.size(); ++i) {
            if (!vis[i]) continue;
            const Face& f = faces[i];
            owner[{f.a, f.b}] = i;
            owner[{f.b, f.c}] = i;
            owner[{f.c, f.a}] = i;
        }
        std::vector<std::pair<uint32_t, uint32_t>> horizon;
        for (const auto& kv : owner)
            if (owner.find({kv.first.second, kv.first.first}) == owner.end()) horizon.push_back(kv.first);
        std::vector<Face> next;
        for (size_t i = 0; i < faces.size(); ++i)
            if (!vis[i]) next.push_back(faces[i]);
        for (const auto& e : horizon) {
            Face f;
            f.a = e.first; f.b = e.second; f.c = pt;
            hull_detail::face_plane(f, cloud);
            next.push_back(f);
        }
        faces = std::move(next);
        used[pt] = true;
    }
    // coplanar merge
    std::vector<bool> merged(faces.size(), false);
    std::vector<std::vector<uint32_t>> polys;
    std::vector<V3<float>> pnorm;
    for (size_t i = 0; i < faces.size(); ++i) {
        if (merged[i]) continue;
        std::vector<size_t> group{i};
        merged[i] = true;
        for (size_t j = i + 1; j < faces.size(); ++j) {
            if (merged[j]) continue;
            if (dot(faces[i].n, faces[j].n) > 1.0f - tol &&
                std::abs(faces[i].d - faces[j].d) < std::max(tol, 1.0f) * tol * 100.0f) {
                group.push_back(j);
                merged[j] = true;
            }
        }
        V3<float> n = faces[i].n;
        float d = faces[i].d;
        std::vector<uint32_t> verts;
        for (size_t g : group)
            for (uint32_t v : {faces[g].a, faces[g].b, faces[g].c})
                if (std::find(verts.begin(), verts.end(), v) == verts.end()) verts.push_back(v);
        V3<float> u = std::abs(n.x) < 0.9f ? normalize(cross(n, V3<float>(1, 0, 0))) : normalize(cross(n, V3<float>(0, 1, 0)));
        V3<float> w = cross(n, u);
        V3<float> origin = n * d;
        std::vector<hull_detail::V2> proj(verts.size());
        for (size_t k = 0; k < verts.size(); ++k) {
            V3<float> rel = cloud[verts[k]] - origin;
            proj[k] = hull_detail::V2{dot(rel, u), dot(rel, w)};
        }
        std::vector<uint32_t> h2 = hull_detail::hull_2d(proj);
        if (h2.size() < 3) continue;
        std::vector<uint32_t> poly(h2.size());
        for (size_t k = 0; k < h2.size(); ++k) poly[k] = verts[h2[k]];
        polys.push_back(poly);
        pnorm.push_back(n);
    }
    std::vector<uint32_t> remap(cloud.size(), UINT32_MAX);
    for (const auto& poly : polys)
        for (uint32_t v : poly)
            if (remap[v] == UINT32_MAX) { remap[v] = uint32_t(out.v.size()); out.v.push_back(cloud[v]); }
    out.fs.push_back(0);
    std::map<std::pair<uint32_t, uint32_t>, bool> seen;
    for (size_t i = 0; i < polys.size(); ++i) {
        const auto& poly = polys[i];
        out.pn.push_back(pnorm[i]);
        out.pd.push_back(dot(pnorm[i], cloud[poly[0]]));
        for (size_t k = 0; k < poly.size(); ++k) {
            uint32_t ra = remap[poly[k]], rb = remap[poly[(k + 1) % poly.size()]];
            out.fv.push_back(ra);
            auto key = std::make_pair(std::min(ra, rb), std::max(ra, rb));
            if (!seen[key]) { seen[key] = true; out.e.push_back(key); }
        }
        out.fs.push_back(uint32_t(out.fv.size()));
    }
    if (out.faces() < 4 || out.v.size() < 4) return false;
    V3<float> ref(0.0f);
    for (auto& p : out.v) ref += p;
    ref = ref / float(out.v.size());
    V3<float> wsum(0.0f);
    float tv = 0.0f;
    for (size_t i = 0; i < polys.size(); ++i) {
        uint32_t s = out.fs[i], e = out.fs[i + 1];
        V3<float> v0 = out.v[out.fv[s]];
        for (uint32_t k = s + 1; k + 1 < e; ++k) {
            V3<float> v1 = out.v[out.fv[k]], v2 = out.v[out.fv[k + 1]];
            float vol6 = dot(v0 - ref, cross(v1 - ref, v2 - ref));
            V3<float> tc = (ref + v0 + v1 + v2) * 0.25f;
            tv += vol6;
            wsum += tc * vol6;
        }
    }
    out.volume = std::abs(tv) / 6.0f;
    out.centroid = tv != 0.0f ? wsum / tv : ref;
    out.lb = AABB<float>{};
    for (auto& p : out.v) { out.lb.mn = vmin(out.lb.mn, p); out.lb.mx = vmax(out.lb.mx, p); }
    return true;
}

// ---- narrowphase (collision/sat_hull.h) ----

template <class T> struct PosedHull {  // sat_hull.h:47-60
    const Hull<T>* h = nullptr;
    V3<T> pos{T(0)};
    Q<T> rot;
    T scale = T(1);
    V3<T> vtx(uint32_t i) const { return pos + qrot(rot, h->v[i] * scale); }
    V3<T> fnormal(size_t f) const { return qrot(rot, h->pn[f]); }
};

template <class T> struct HullInterval { T lo, hi; };
template <class T> inline HullInterval<T> hproject(const std::vector<V3<T>>& pts, const V3<T>& axis) {  // sat_hull.h:84-92
    HullInterval<T> r{std::numeric_limits<T>::max(), -std::numeric_limits<T>::max()};
    for (const V3<T>& p : pts) {
        T d = dot(p, axis);
        r.lo = std::min(r.lo, d);
        r.hi = std::max(r.hi, d);
    }
    return r;
}
template <class T> inline std::vector<V3<T>> hull_world_vertices(const PosedHull<T>& h) {  // sat_hull.h:77-81
    std::vector<V3<T>> out(h.h->v.size());
    for (size_t i = 0; i < out.size(); ++i) out[i] = h.vtx(uint32_t(i));
    return out;
}
template <class T> inline bool face_overlap(const std::vector<V3<T>>& self, const std::vector<V3<T>>& other, const V3<T>& n,
                                            T& overlap) {  // sat_hull.h:117-123
    const T d = hproject(self, n).hi;
    const T omin = hproject(other, n).lo;
    overlap = d - omin;
    return overlap >= T(0);
}
constexpr int k_hull_clip_cap = 160;
template <class T> inline void clip_polygon_n(const V3<T>* in, int nin, const Plane<T>& pl, V3<T>* out, int& nout) {  // sat_hull.h:129-145
    nout = 0;
    if (nin == 0) return;
    for (int i = 0; i < nin; ++i) {
        const V3<T>& cur = in[i];
        const V3<T>& prev = in[(i - 1 + nin) % nin];
        T dc = dot(pl.n, cur) - pl.d;
        T dp = dot(pl.n, prev) - pl.d;
        bool ci = dc <= T(0), pi = dp <= T(0);
        if (ci != pi) {
            T t = dp / (dp - dc);
            if (nout < k_hull_clip_cap) out[nout++] = prev + t * (cur - prev);
        }
        if (ci && nout < k_hull_clip_cap) out[nout++] = cur;
    }
}
template <class T> inline V3<T> outward_edge_normal(const V3<T>& edge, const V3<T>& fn) {  // sat_hull.h:157-161
    V3<T> n = cross(edge, fn);
    T len = length(n);
    return len > T(1e-12f) ? n / len : fn;
}
template <class T> inline int hull_face_side_planes(const PosedHull<T>& h, size_t face, Plane<T>* pl, int cap) {  // sat_hull.h:165-178
    uint32_t s = h.h->fs[face], e = h.h->fs[face + 1];
    int n = int(e - s);
    V3<T> normal = h.fnormal(face);
    int count = 0;
    for (int i = 0; i < n && count < cap; ++i) {
        V3<T> v0 = h.vtx(h.h->fv[s + i]);
        V3<T> v1 = h.vtx(h.h->fv[s + (i + 1) % n]);
        V3<T> out = outward_edge_normal(v1 - v0, normal);
        pl[count++] = Plane<T>{out, dot(out, v0)};
    }
    return count;
}
template <class T> inline int hull_face_polygon(const PosedHull<T>& h, size_t face, V3<T>* out, int cap) {  // sat_hull.h:181-187
    uint32_t s = h.h->fs[face], e = h.h->fs[face + 1];
    int n = std::min(int(e - s), cap);
    for (int i = 0; i < n; ++i) out[i] = h.vtx(h.h->fv[s + i]);
    return n;
}

template <class T> struct HullSat {  // sat_hull.h:203-209
    T pen = 0;
    V3<T> n{T(0), T(0), T(1)};
    int kind = 0;  // 0 face of A, 1 face of B, 2 edge-edge
    int fa = -1, fb = -1, ea = -1, eb = -1;
};

template <class T> bool sat_hull_hull(const PosedHull<T>& a, const PosedHull<T>& b, HullSat<T>& r) {  // sat_hull.h:223-283
    std::vector<V3<T>> pa = hull_world_vertices(a), pb = hull_world_vertices(b);
    T best = std::numeric_limits<T>::max();
    bool found = false;
    const T bias = T(0.01f);
    auto face = [&](const std::vector<V3<T>>& self, const std::vector<V3<T>>& other, const V3<T>& nraw, bool self_a,
                    int kind, int fa, int fb) -> bool {
        T len2 = dot(nraw, nraw);
        if (len2 < T(1e-10f)) return true;
        V3<T> n = nraw / std::sqrt(len2);
        T ov = 0;
        if (!face_overlap(self, other, n, ov)) return false;
        if (ov < best) {
            best = ov;
            r.pen = ov; r.n = self_a ? n : -n; r.kind = kind; r.fa = fa; r.fb = fb; r.ea = -1; r.eb = -1;
            found = true;
        }
        return true;
    };
    for (size_t f = 0; f < a.h->faces(); ++f)
        if (!face(pa, pb, a.fnormal(f), true, 0, int(f), -1)) return false;
    for (size_t f = 0; f < b.h->faces(); ++f)
        if (!face(pb, pa, b.fnormal(f), false, 1, -1, int(f))) return false;
    for (size_t i = 0; i < a.h->e.size(); ++i) {
        V3<T> ed = a.vtx(a.h->e[i].second) - a.vtx(a.h->e[i].first);
        for (size_t j = 0; j < b.h->e.size(); ++j) {
            V3<T> fd = b.vtx(b.h->e[j].second) - b.vtx(b.h->e[j].first);
            V3<T> ar = cross(ed, fd);
            T len2 = dot(ar, ar);
            if (len2 < T(1e-10f)) continue;
            V3<T> axis = ar / std::sqrt(len2);
            HullInterval<T> ia = hproject(pa, axis), ib = hproject(pb, axis);
            T ov = std::min(ia.hi - ib.lo, ib.hi - ia.lo);
            if (ov < T(0)) return false;
            T score = ov + bias;
            if (score < best) {
                best = score;
                r.pen = ov;
                r.n = (ia.lo + ia.hi) >= (ib.lo + ib.hi) ? -axis : axis;
                r.kind = 2; r.fa = -1; r.fb = -1; r.ea = int(i); r.eb = int(j);
                found = true;
            }
        }
    }
    return found;
}

// the contact candidates of a clipped polygon below a reference plane (sat_hull.h:355-375):
// kept when deeper than -epsilon, at most 4, deepest first (std::sort on at most 16 candidates
// is libstdc++'s insertion sort, which keeps equal depths in clip order)
template <class T> struct HullCand { V3<T> pos; T pen; };
template <class T> inline int hull_keep(const V3<T>* poly, int count, const V3<T>& rn, T ref_d, HullCand<T>* cand) {
    int kept = 0;
    for (int i = 0; i < count; ++i) {
        T pen = -(dot(rn, poly[i]) - ref_d);
        if (pen > -k_epsilon<T>()) { cand[kept].pos = poly[i]; cand[kept].pen = std::max(pen, T(0)); ++kept; }
    }
    if (kept > 4) {
        std::sort(cand, cand + kept, [](const HullCand<T>& x, const HullCand<T>& y) { return x.pen > y.pen; });
        kept = 4;
    }
    return kept;
}

template <class T> bool hull_hull_contacts(const PosedHull<T>& a, const PosedHull<T>& b, Manifold<T>& out) {  // sat_hull.h:297-380
    HullSat<T> r;
    if (!sat_hull_hull(a, b, r)) return false;
    out = Manifold<T>{};
    out.n = r.n;
    if (r.kind == 2) {
        V3<T> ea0 = a.vtx(a.h->e[r.ea].first), ea1 = a.vtx(a.h->e[r.ea].second);
        V3<T> eb0 = b.vtx(b.h->e[r.eb].first), eb1 = b.vtx(b.h->e[r.eb].second);
        V3<T> c1, c2;
        closest_points_seg_seg(ea0, ea1, eb0, eb1, c1, c2);
        out.add_point(T(0.5f) * (c1 + c2), r.pen, 0x50000000u | (uint32_t(r.ea) << 16) | uint32_t(r.eb));
        return true;
    }
    bool ref_a = r.kind == 0;
    const PosedHull<T>& ref = ref_a ? a : b;
    const PosedHull<T>& inc = ref_a ? b : a;
    size_t rf = ref_a ? size_t(r.fa) : size_t(r.fb);
    V3<T> rn = ref.fnormal(rf);
    size_t inf = 0;
    T most = std::numeric_limits<T>::max();
    for (size_t f = 0; f < inc.h->faces(); ++f) {
        T dp = dot(inc.fnormal(f), rn);
        if (dp < most) { most = dp; inf = f; }
    }
    std::vector<V3<T>> poly(k_hull_clip_cap), poly2(k_hull_clip_cap);
    int count = hull_face_polygon(inc, inf, poly.data(), k_hull_clip_cap);
    std::vector<Plane<T>> planes(k_hull_clip_cap);
    int np = hull_face_side_planes(ref, rf, planes.data(), k_hull_clip_cap);
    for (int p = 0; p < np && count > 0; ++p) {
        int nc = 0;
        clip_polygon_n(poly.data(), count, planes[p], poly2.data(), nc);
        for (int i = 0; i < nc; ++i) poly[i] = poly2[i];
        count = nc;
    }
    if (count == 0) return false;
    T ref_d = dot(rn, ref.vtx(ref.h->fv[ref.h->fs[rf]]));
    std::vector<HullCand<T>> cand(k_hull_clip_cap);
    int kept = hull_keep(poly.data(), count, rn, ref_d, cand.data());
    if (kept == 0) return false;
    for (int i = 0; i < kept; ++i)
        out.add_point(cand[i].pos, cand[i].pen, 0x40000000u | (uint32_t(rf) << 16) | (uint32_t(inf) << 8) | uint32_t(i & 0xFF));
    return out.count > 0;
}

template <class T> bool hull_box_contacts(const PosedHull<T>& a, const OBB<T>& b, Manifold<T>& out) {  // sat_hull.h:390-578
    std::vector<V3<T>> pa = hull_world_vertices(a);
    V3<T> bax[3];
    obb_axes(b, bax);
    std::vector<V3<T>> pb;
    for (int sx = -1; sx <= 1; sx += 2)
        for (int sy = -1; sy <= 1; sy += 2)
            for (int sz = -1; sz <= 1; sz += 2)
                pb.push_back(b.c + (T(sx) * b.he.x) * bax[0] + (T(sy) * b.he.y) * bax[1] + (T(sz) * b.he.z) * bax[2]);
    T best = std::numeric_limits<T>::max();
    bool found = false;
    const T bias = T(0.01f);
    int kind = 0, fa = -1, ba = -1, ea = -1, eb = -1;
    T bpen = 0;
    V3<T> bn(T(0), T(0), T(1));
    auto consider = [&](const V3<T>& ar, T bs, int k2, int f_a, int b_a, int e_a, int e_b) -> bool {
        T len2 = dot(ar, ar);
        if (len2 < T(1e-10f)) return true;
        V3<T> axis = ar / std::sqrt(len2);
        HullInterval<T> ia = hproject(pa, axis), ib = hproject(pb, axis);
        T ov = std::min(ia.hi - ib.lo, ib.hi - ia.lo);
        if (ov < T(0)) return false;
        T score = ov + bs;
        if (score < best) {
            best = score; bpen = ov;
            bn = (ia.lo + ia.hi) >= (ib.lo + ib.hi) ? -axis : axis;
            kind = k2; fa = f_a; ba = b_a; ea = e_a; eb = e_b;
            found = true;
        }
        return true;
    };
    for (size_t f = 0; f < a.h->faces(); ++f) {
        V3<T> n = a.fnormal(f);
        T ov = 0;
        if (!face_overlap(pa, pb, n, ov)) return false;
        if (ov < best) {
            best = ov; bpen = ov; bn = n; kind = 0; fa = int(f); ba = -1; ea = -1; eb = -1;
            found = true;
        }
    }
    for (int i = 0; i < 3; ++i)
        if (!consider(bax[i], T(0), 1, -1, i, -1, -1)) return false;
    for (size_t i = 0; i < a.h->e.size(); ++i) {
        V3<T> ed = a.vtx(a.h->e[i].second) - a.vtx(a.h->e[i].first);
        for (int j = 0; j < 3; ++j)
            if (!consider(cross(ed, bax[j]), bias, 2, -1, -1, int(i), j)) return false;
    }
    if (!found) return false;
    out = Manifold<T>{};
    out.n = bn;
    if (kind == 2) {
        V3<T> ea0 = a.vtx(a.h->e[ea].first), ea1 = a.vtx(a.h->e[ea].second);
        V3<T> mid = T(0.5f) * (ea0 + ea1);
        V3<T> d = mid - b.c;
        int aj = (eb + 1) % 3, ak = (eb + 2) % 3;
        T sj = dot(d, bax[aj]) >= T(0) ? T(1) : T(-1);
        T sk = dot(d, bax[ak]) >= T(0) ? T(1) : T(-1);
        V3<T> ec = b.c + bax[aj] * (sj * b.he[aj]) + bax[ak] * (sk * b.he[ak]);
        V3<T> eb0 = ec - bax[eb] * b.he[eb], eb1 = ec + bax[eb] * b.he[eb];
        V3<T> c1, c2;
        closest_points_seg_seg(ea0, ea1, eb0, eb1, c1, c2);
        out.add_point(T(0.5f) * (c1 + c2), bpen, 0x60000000u | (uint32_t(ea) << 8) | uint32_t(eb));
        return true;
    }
    std::vector<V3<T>> poly(k_hull_clip_cap), poly2(k_hull_clip_cap);
    std::vector<HullCand<T>> cand(k_hull_clip_cap);
    if (kind == 0) {
        V3<T> rn = a.fnormal(size_t(fa));
        int inc_axis = 0;
        T inc_sign = T(1);
        T most = std::numeric_limits<T>::max();
        for (int ax = 0; ax < 3; ++ax)
            for (int si = 0; si < 2; ++si) {
                T s = si == 0 ? T(1) : T(-1);
                T dp = dot(bax[ax] * s, rn);
                if (dp < most) { most = dp; inc_axis = ax; inc_sign = s; }
            }
        V3<T> ic[4], inn;
        obb_face(b, inc_axis, inc_sign, ic, inn);
        int count = 4;
        for (int i = 0; i < 4; ++i) poly[i] = ic[i];
        std::vector<Plane<T>> planes(k_hull_clip_cap);
        int np = hull_face_side_planes(a, size_t(fa), planes.data(), k_hull_clip_cap);
        for (int p = 0; p < np && count > 0; ++p) {
            int nc = 0;
            clip_polygon_n(poly.data(), count, planes[p], poly2.data(), nc);
            for (int i = 0; i < nc; ++i) poly[i] = poly2[i];
            count = nc;
        }
        if (count == 0) return false;
        T ref_d = dot(rn, a.vtx(a.h->fv[a.h->fs[fa]]));
        int kept = hull_keep(poly.data(), count, rn, ref_d, cand.data());
        if (kept == 0) return false;
        for (int i = 0; i < kept; ++i)
            out.add_point(cand[i].pos, cand[i].pen, 0x40000000u | (uint32_t(fa) << 16) | uint32_t(i & 0xFF));
        return out.count > 0;
    }
    V3<T> desired = -bn;
    T ref_sign = dot(bax[ba], desired) >= T(0) ? T(1) : T(-1);
    V3<T> rc[4], rn;
    obb_face(b, ba, ref_sign, rc, rn);
    size_t inf = 0;
    T most = std::numeric_limits<T>::max();
    for (size_t f = 0; f < a.h->faces(); ++f) {
        T dp = dot(a.fnormal(f), rn);
        if (dp < most) { most = dp; inf = f; }
    }
    int count = hull_face_polygon(a, inf, poly.data(), k_hull_clip_cap);
    Plane<T> planes[4];
    obb_side_planes(b, ba, planes);
    for (int p = 0; p < 4 && count > 0; ++p) {
        int nc = 0;
        clip_polygon_n(poly.data(), count, planes[p], poly2.data(), nc);
        for (int i = 0; i < nc; ++i) poly[i] = poly2[i];
        count = nc;
    }
    if (count == 0) return false;
    T ref_d = dot(rn, rc[0]);
    int kept = hull_keep(poly.data(), count, rn, ref_d, cand.data());
    if (kept == 0) return false;
    for (int i = 0; i < kept; ++i)
        out.add_point(cand[i].pos, cand[i].pen, 0x70000000u | (uint32_t(ba) << 16) | (uint32_t(inf) << 8) | uint32_t(i & 0xFF));
    return out.count > 0;
}

template <class T> V3<T> closest_on_convex_polygon(const V3<T>* poly, int n, const V3<T>& normal, const V3<T>& point) {  // sat_hull.h:594-625
    const T pd = dot(normal, poly[0]);
    const V3<T> proj = point - normal * (dot(normal, point) - pd);
    bool inside = true;
    for (int i = 0; i < n && inside; ++i) {
        const V3<T> a = poly[i], b = poly[(i + 1) % n];
        const V3<T> out = outward_edge_normal(b - a, normal);
        if (dot(out, proj - a) > T(1e-6f)) inside = false;
    }
    if (inside) return proj;
    V3<T> best = poly[0];
    T bd = std::numeric_limits<T>::max();
    for (int i = 0; i < n; ++i) {
        V3<T> a = poly[i], b = poly[(i + 1) % n];
        V3<T> ab = b - a;
        T len2 = dot(ab, ab);
        T t = len2 > T(1e-12f) ? sclamp(dot(proj - a, ab) / len2, T(0), T(1)) : T(0);
        V3<T> p = a + ab * t;
        T d2 = dot(proj - p, proj - p);
        if (d2 < bd) { bd = d2; best = p; }
    }
    return best;
}

template <class T> bool hull_vs_sphere(const PosedHull<T>& a, const Sph<T>& b, V3<T>& normal, T& pen, V3<T>& point) {  // sat_hull.h:647-687
    size_t bf = 0;
    T bd = -std::numeric_limits<T>::max();
    bool any_pos = false;
    for (size_t f = 0; f < a.h->faces(); ++f) {
        V3<T> n = a.fnormal(f);
        T d = dot(n, a.vtx(a.h->fv[a.h->fs[f]]));
        T dist = dot(n, b.c) - d;
        if (dist > T(0)) any_pos = true;
        if (dist > bd) { bd = dist; bf = f; }
    }
    V3<T> n = a.fnormal(bf);
    std::vector<V3<T>> poly(k_hull_clip_cap);
    int count = hull_face_polygon(a, bf, poly.data(), k_hull_clip_cap);
    if (!any_pos) {
        point = closest_on_convex_polygon(poly.data(), count, n, b.c);
        V3<T> diff = point - b.c;
        T dist = length(diff);
        normal = dist > k_epsilon<T>() ? diff / dist : n;
        pen = b.r + (-bd);
        return true;
    }
    point = closest_on_convex_polygon(poly.data(), count, n, b.c);
    V3<T> diff = b.c - point;
    T dist = length(diff);
    if (dist > b.r) return false;
    normal = dist > k_epsilon<T>() ? diff / dist : n;
    pen = b.r - dist;
    return true;
}

template <class T> bool hull_vs_capsule(const PosedHull<T>& a, const Cap<T>& b, Manifold<T>& out) {  // sat_hull.h:698-722
    out = Manifold<T>{};
    V3<T> samples[3] = {b.a, b.b, T(0.5f) * (b.a + b.b)};
    V3<T> bn(T(0), T(0), T(1));
    T bp = 0;
    bool hit = false;
    for (int i = 0; i < 3; ++i) {
        Sph<T> probe{samples[i], b.r};
        V3<T> n, p;
        T pen;
        if (hull_vs_sphere(a, probe, n, pen, p)) {
            out.add_point(p, pen, 0x80000000u | uint32_t(i));
            if (pen > bp) { bp = pen; bn = n; }
            hit = true;
        }
    }
    if (!hit) return false;
    out.n = bn;
    return true;
}

// ---- the hull shape in the world: pose, bounds, CCD extent and the narrowphase dispatch ----

template <class T> inline PosedHull<T> world_posed_hull(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot) {  // sat_hull.h:65-72
    PosedHull<T> h;
    h.h = static_cast<const Hull<T>*>(s.hull);
    h.pos = pos + qrot(rot, s.lc);
    h.rot = qmul(rot, s.lr);
    h.scale = s.hull_scale;
    return h;
}
// shape.h:222-225 + convex_hull.h:83-87: the local bounds' 8 corners through mat4(mat3_cast(rot) * scale)
template <class T> AABB<T> hull_world_bounds(const Shape<T>& s, const V3<T>& pos, const Q<T>& rot) {
    const Hull<T>* h = static_cast<const Hull<T>*>(s.hull);
    if (!h) return AABB<T>{};
    const V3<T> c = pos + qrot(rot, s.lc);
    M3<T> m = mat3_cast(qmul(rot, s.lr));
    for (int k = 0; k < 3; ++k) m.c[k] = m.c[k] * s.hull_scale;
    AABB<T> out;
    for (int i = 0; i < 8; ++i) {
        V3<T> corner((i & 1) ? h->lb.mx.x : h->lb.mn.x, (i & 2) ? h->lb.mx.y : h->lb.mn.y, (i & 4) ? h->lb.mx.z : h->lb.mn.z);
        V3<T> wpt = affine(m, c, corner);
        out.mn = vmin(out.mn, wpt);
        out.mx = vmax(out.mx, wpt);
    }
    return out;
}
template <class T> T hull_min_extent(const Shape<T>& s) {  // world.h:1417-1420
    const Hull<T>* h = static_cast<const Hull<T>*>(s.hull);
    if (!h) return std::numeric_limits<T>::max();
    const V3<T> half = (h->lb.mx - h->lb.mn) * T(0.5f) * s.hull_scale;
    return std::min(std::min(half.x, half.y), half.z);
}
// narrowphase.h:271-313: the hull branches of generate_contacts (no speculative contacts);
// false when neither shape is a hull
template <class T> bool hull_pair_contacts(const Shape<T>& sa, const V3<T>& pa, const Q<T>& ra, const Shape<T>& sb,
                                           const V3<T>& pb, const Q<T>& rb, Manifold<T>& out, bool& hit) {
    const int ta = sa.type, tb = sb.type;
    if (ta != PXG_CONVEX_HULL && tb != PXG_CONVEX_HULL) return false;
    if (!sa.hull && ta == PXG_CONVEX_HULL) return false;
    if (!sb.hull && tb == PXG_CONVEX_HULL) return false;
    V3<T> n, p;
    T pen;
    hit = false;
    if (ta == PXG_CONVEX_HULL && tb == PXG_CONVEX_HULL) {
        hit = hull_hull_contacts(world_posed_hull(sa, pa, ra), world_posed_hull(sb, pb, rb), out);
    } else if (ta == PXG_CONVEX_HULL && tb == PXG_BOX) {
        hit = hull_box_contacts(world_posed_hull(sa, pa, ra), world_obb(sb, pb, rb), out);
    } else if (ta == PXG_BOX && tb == PXG_CONVEX_HULL) {
        hit = hull_box_contacts(world_posed_hull(sb, pb, rb), world_obb(sa, pa, ra), out);
        if (hit) out.n = -out.n;
    } else if (ta == PXG_CONVEX_HULL && tb == PXG_SPHERE) {
        if (hull_vs_sphere(world_posed_hull(sa, pa, ra), world_sphere(sb, pb, rb), n, pen, p)) {
            out.n = n; out.add_point(p, pen, 0); hit = true;
        }
    } else if (ta == PXG_SPHERE && tb == PXG_CONVEX_HULL) {
        if (hull_vs_sphere(world_posed_hull(sb, pb, rb), world_sphere(sa, pa, ra), n, pen, p)) {
            out.n = -n; out.add_point(p, pen, 0); hit = true;
        }
    } else if (ta == PXG_CONVEX_HULL && tb == PXG_CAPSULE) {
        hit = hull_vs_capsule(world_posed_hull(sa, pa, ra), world_capsule(sb, pb, rb), out);
    } else if (ta == PXG_CAPSULE && tb == PXG_CONVEX_HULL) {
        hit = hull_vs_capsule(world_posed_hull(sb, pb, rb), world_capsule(sa, pa, ra), out);
        if (hit) out.n = -out.n;
    } else {
        return false;
    }
    return true;
}

}  // namespace pxo
