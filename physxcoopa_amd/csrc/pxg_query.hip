// Scene raycasts against one world (query/queries.h, geometry/ray.h, world.h:766-847).
//
// The reference walks its two AABB trees with the ray's slab test against the proxies' fat AABBs
// and tests every leaf it reaches exactly. A leaf's box lies inside its ancestors' and the slab
// test is monotonic in the box, so the set of shapes tested is exactly the live shapes whose fat
// AABB the slab test accepts; the GPU tests that set directly, culled by a two-level layout built
// per query batch: a prep pass turns every shape into its world geometry (sphere centre and
// radius, box centre, half extents and orientation, capsule end points and radius) from the
// bodies' current poses; shapes and rays are sorted by the Morton code of their fat-box centre /
// origin; consecutive 256-shape tiles of the sorted shapes get their bounding box. One thread owns
// one ray, a block's 256 rays are neighbours in Morton order, and a block loads a tile into LDS
// only when the slab test accepts the tile's box for one of its rays (a shape's fat box lies in
// its tile's box, so nothing the reference tests is skipped). The closest hit keeps the nearest
// distance, equal distances going to the lower shape slot.
#include "pxg_state.h"
#include "../../include/pxg.h"
#include "pxg_device.h"

#include <rocprim/device/device_radix_sort.hpp>

using namespace pxd;

namespace pxg {
namespace {

constexpr uint32_t kQueryTile = 256u;  // shapes per LDS tile = rays per block
constexpr uint32_t QF_LIVE = 1u, QF_TRIGGER = 2u;

// glm::inverse(quat): conjugate / dot(q, q)
__device__ __forceinline__ quat qinverse_q(quat q) {
    float d = (q.x * q.x + q.y * q.y) + (q.z * q.z + q.w * q.w);
    quat r;
    r.w = q.w / d; r.x = -q.x / d; r.y = -q.y / d; r.z = -q.z / d;
    return r;
}
// std::max / std::min over an initializer list: the first largest / smallest element
__device__ __forceinline__ float max4(float a, float b, float c, float d) {
    float m = a;
    if (m < b) m = b;
    if (m < c) m = c;
    if (m < d) m = d;
    return m;
}
__device__ __forceinline__ float min4(float a, float b, float c, float d) {
    float m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    if (d < m) m = d;
    return m;
}
__device__ __forceinline__ v3 vmul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }

// order-preserving float <-> uint encoding (for atomic min / max of floats)
__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u); }
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
// Morton code of p in the box [lo, hi] (10 bits per axis, clamped)
__device__ __forceinline__ uint32_t morton(v3 p, v3 lo, v3 hi) {
    uint32_t q[3];
    const float pc[3] = {p.x, p.y, p.z}, lc[3] = {lo.x, lo.y, lo.z}, hc[3] = {hi.x, hi.y, hi.z};
    for (int k = 0; k < 3; ++k) {
        const float ext = hc[k] - lc[k];
        float u = ext > 0.0f ? (pc[k] - lc[k]) / ext : 0.0f;
        u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
        q[k] = uint32_t(u * 1023.0f);
    }
    return spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
}

struct QRay {
    v3 o, d;
    float maxd;
};

// Ray::intersect (geometry/ray.h:47-60)
__device__ __forceinline__ bool ray_aabb(const QRay& r, v3 mn, v3 mx) {
    const float e = 1e-9f;
    v3 inv = mk(1.0f / (r.d.x + e), 1.0f / (r.d.y + e), 1.0f / (r.d.z + e));
    v3 t0 = vmul(mn - r.o, inv), t1 = vmul(mx - r.o, inv);
    v3 tmin = vmin(t0, t1), tmax = vmax(t0, t1);
    float t_near = max4(tmin.x, tmin.y, tmin.z, 0.0f);
    float t_far = min4(tmax.x, tmax.y, tmax.z, r.maxd);
    return !(t_far < t_near);
}
// ray_vs_sphere (geometry/ray.h:62-78)
__device__ __forceinline__ bool ray_sphere(const QRay& r, v3 c, float rad, float maxd, float& t, v3& n) {
    v3 oc = r.o - c;
    float b = dot(oc, r.d);
    float cc = dot(oc, oc) - rad * rad;
    float disc = b * b - cc;
    if (disc < 0.0f) return false;
    float s = sqrtf(disc);
    float t0 = -b - s, t1 = -b + s;
    float hit = t0 >= 0.0f ? t0 : t1;
    if (hit < 0.0f || hit > maxd) return false;
    t = hit;
    n = normalize((r.o + r.d * t) - c);
    return true;
}
// ray_vs_obb (geometry/ray.h:101-125)
__device__ __forceinline__ bool ray_obb(const QRay& r, v3 c, v3 he, quat q, float& t, v3& n) {
    quat inv = qinverse_q(q);
    v3 lo = qrot(inv, r.o - c);
    v3 ld = qrot(inv, r.d);
    const float e = 1e-9f;
    v3 invd = mk(1.0f / (ld.x + e), 1.0f / (ld.y + e), 1.0f / (ld.z + e));
    v3 t0 = vmul(-he - lo, invd), t1 = vmul(he - lo, invd);
    v3 tmin = vmin(t0, t1), tmax = vmax(t0, t1);
    float t_near = max4(tmin.x, tmin.y, tmin.z, 0.0f);
    float t_far = min4(tmax.x, tmax.y, tmax.z, r.maxd);
    if (t_far < t_near) return false;
    int axis = 0;
    if (tmin.y > tmin.x) axis = 1;
    float ta = axis == 0 ? tmin.x : tmin.y;
    if (tmin.z > ta) axis = 2;
    float lda = axis == 0 ? ld.x : (axis == 1 ? ld.y : ld.z);
    float sgn = lda >= 0.0f ? -1.0f : 1.0f;
    v3 ln = mk(axis == 0 ? sgn : 0.0f, axis == 1 ? sgn : 0.0f, axis == 2 ? sgn : 0.0f);
    t = t_near;
    n = qrot(q, ln);
    return true;
}
// ray_vs_capsule (geometry/ray.h:138-194)
__device__ __forceinline__ bool ray_capsule(const QRay& r, v3 a, v3 b, float rad, float& t, v3& n) {
    v3 axis = b - a;
    float axis_len2 = dot(axis, axis);
    bool found = false;
    float best_t = r.maxd;
    v3 best_n = mk(0.0f, 0.0f, 0.0f);
    if (axis_len2 > 1e-6f) {
        v3 axis_n = axis / sqrtf(axis_len2);
        v3 oc = r.o - a;
        v3 d_perp = r.d - axis_n * dot(r.d, axis_n);
        v3 oc_perp = oc - axis_n * dot(oc, axis_n);
        float qa = dot(d_perp, d_perp);
        if (qa > 1e-6f) {
            float qb = dot(d_perp, oc_perp);
            float qc = dot(oc_perp, oc_perp) - rad * rad;
            float disc = qb * qb - qa * qc;
            if (disc >= 0.0f) {
                float s = sqrtf(disc);
                float cands[2] = {(-qb - s) / qa, (-qb + s) / qa};
                for (int k = 0; k < 2; ++k) {
                    float cand = cands[k];
                    if (cand < 0.0f || cand >= best_t) continue;
                    float along = dot((r.o + r.d * cand) - a, axis_n);
                    if (along < 0.0f || along > sqrtf(axis_len2)) continue;
                    v3 hp = r.o + r.d * cand;
                    v3 cl = a + axis_n * along;
                    best_t = cand;
                    best_n = normalize(hp - cl);
                    found = true;
                    break;
                }
            }
        }
    }
    for (int k = 0; k < 2; ++k) {
        float ct;
        v3 cn;
        if (ray_sphere(r, k == 0 ? a : b, rad, best_t, ct, cn)) {
            best_t = ct;
            best_n = cn;
            found = true;
        }
    }
    if (!found) return false;
    t = best_t;
    n = best_n;
    return true;
}

// One shape's query record: fat AABB, world geometry (g0, g1, g2 by type: sphere (c, r);
// box (c, -), (he, -), q; capsule (a, r), (b, -)), and (owner, type, flags, layer).
struct QShape {
    float4 fmin, fmax, g0, g1, g2;
    uint4 info;
};

__device__ __forceinline__ bool shape_ray(const QRay& r, const QShape& s, float& t, v3& n) {
    if (s.info.y == SH_SPHERE) return ray_sphere(r, ld3(s.g0), s.g0.w, r.maxd, t, n);
    if (s.info.y == SH_BOX) return ray_obb(r, ld3(s.g0), ld3(s.g1), ldq(s.g2), t, n);
    if (s.info.y == SH_CAPSULE) return ray_capsule(r, ld3(s.g0), ld3(s.g1), s.g0.w, t, n);
    return false;
}
// QueryFilter::accepts without the predicate (query/queries.h:92-98), for a live proxy
__device__ __forceinline__ bool accepts(const uint4 info, const QueryFilterDev& f) {
    if (!(info.z & QF_LIVE)) return false;
    if (!((f.layer_mask >> (info.w & 31u)) & 1u)) return false;
    if (!f.include_triggers && (info.z & QF_TRIGGER)) return false;
    if (f.ignore_body != 0xFFFFFFFFu && info.x == f.ignore_body) return false;
    return true;
}

// world_sphere / world_obb / world_capsule (collision/shape.h:160-188) from the current poses
__global__ void k_query_prep(DevState d, float4* geom, uint4* qinfo, uint32_t* bounds6) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.ns) return;
    const uint4 info = d.s_info[i];
    ShapeGeom s;
    s.type = int(info.y & 15u);
    s.caxis = int((info.y >> 8) & 3u);
    const float4 lc = d.s_lc[i];
    s.lc = mk(lc.x, lc.y, lc.z);
    s.radius = lc.w;
    s.lr = ldq(d.s_lr[i]);
    const float4 dim = d.s_dim[i];
    s.he = mk(dim.x, dim.y, dim.z);
    s.cr = dim.w;
    s.chh = d.s_mat[i].x;
    const uint32_t owner = info.x;
    // scene queries test sphere, box and capsule shapes; convex hulls are not query targets here
    const bool live = (info.y & SF_ENABLED) && owner < d.nb && (d.b_flags[owner] & BF_ALIVE) && s.type <= SH_CAPSULE;
    float4 g0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), g1 = g0, g2 = g0;
    if (live) {
        const v3 pos = ld3(d.b_pos[owner]);
        const quat rot = ldq(d.b_rot[owner]);
        if (s.type == SH_SPHERE) {
            Sph o = world_sphere(s, pos, rot);
            g0 = st3(o.c, o.r);
        } else if (s.type == SH_BOX) {
            OBB o = world_obb(s, pos, rot);
            g0 = st3(o.c, 0.0f);
            g1 = st3(o.he, 0.0f);
            g2 = stq(o.q);
        } else if (s.type == SH_CAPSULE) {
            Cap o = world_capsule(s, pos, rot);
            g0 = st3(o.a, o.r);
            g1 = st3(o.b, 0.0f);
        }
    }
    geom[3 * i] = g0;
    geom[3 * i + 1] = g1;
    geom[3 * i + 2] = g2;
    qinfo[i] = make_uint4(owner, uint32_t(s.type), (live ? QF_LIVE : 0u) | ((info.y & SF_TRIGGER) ? QF_TRIGGER : 0u),
                          info.z);
    if (live) {
        const v3 c = (ld3(d.s_fmin[i]) + ld3(d.s_fmax[i])) * 0.5f;
        atomicMin(&bounds6[0], f2o(c.x)); atomicMin(&bounds6[1], f2o(c.y)); atomicMin(&bounds6[2], f2o(c.z));
        atomicMax(&bounds6[3], f2o(c.x)); atomicMax(&bounds6[4], f2o(c.y)); atomicMax(&bounds6[5], f2o(c.z));
    }
}
__device__ __forceinline__ void load_bounds(const uint32_t* b6, v3& lo, v3& hi) {
    lo = mk(o2f(b6[0]), o2f(b6[1]), o2f(b6[2]));
    hi = mk(o2f(b6[3]), o2f(b6[4]), o2f(b6[5]));
}
// Morton keys of the shapes' fat-box centres (dead shapes last) and of the rays' origins, with
// identity values for the sort
__global__ void k_query_keys(DevState d, const uint4* qinfo, const uint32_t* bounds6, uint32_t* skey, uint32_t* sidx,
                             const float4* rays, uint32_t n, uint32_t* rkey, uint32_t* ridx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    v3 lo, hi;
    load_bounds(bounds6, lo, hi);
    if (i < d.ns) {
        const bool live = (qinfo[i].z & QF_LIVE) != 0u;
        const v3 c = (ld3(d.s_fmin[i]) + ld3(d.s_fmax[i])) * 0.5f;
        skey[i] = live ? morton(c, lo, hi) : 0x7FFFFFFFu;
        sidx[i] = i;
    }
    if (i < n) {
        // where the ray crosses the scene: the midpoint of its segment inside the centres' box
        // (grown by 1 m); rays that miss the box go last
        const float4 a = rays[2 * i], b = rays[2 * i + 1];
        QRay r;
        r.o = ld3(a);
        r.maxd = a.w;
        r.d = ld3(b);
        const v3 glo = lo - splat(1.0f), ghi = hi + splat(1.0f);
        const v3 inv = mk(1.0f / (r.d.x + 1e-9f), 1.0f / (r.d.y + 1e-9f), 1.0f / (r.d.z + 1e-9f));
        const v3 t0 = vmul(glo - r.o, inv), t1 = vmul(ghi - r.o, inv);
        const v3 tmin = vmin(t0, t1), tmax = vmax(t0, t1);
        const float tn = max4(tmin.x, tmin.y, tmin.z, 0.0f), tf = min4(tmax.x, tmax.y, tmax.z, r.maxd);
        rkey[i] = tf < tn ? 0x3FFFFFFFu : morton(r.o + r.d * (0.5f * (tn + tf)), lo, hi);
        ridx[i] = i;
    }
}
// the sorted shape records (fmin.w = the shape's slot) and each tile's box
__global__ void k_query_gather(DevState d, const float4* geom, const uint4* qinfo, const uint32_t* order, QShape* sorted) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.ns) return;
    const uint32_t j = order[p];
    QShape s;
    s.fmin = d.s_fmin[j];
    s.fmin.w = __uint_as_float(j);
    s.fmax = d.s_fmax[j];
    s.g0 = geom[3 * j];
    s.g1 = geom[3 * j + 1];
    s.g2 = geom[3 * j + 2];
    s.info = qinfo[j];
    sorted[p] = s;
}
__global__ void __launch_bounds__(kQueryTile) k_tile_bounds(DevState d, const QShape* sorted, float4* tile_box) {
    __shared__ float red[6][kQueryTile];
    const uint32_t p = blockIdx.x * kQueryTile + threadIdx.x;
    float v[6] = {3.0e38f, 3.0e38f, 3.0e38f, -3.0e38f, -3.0e38f, -3.0e38f};
    if (p < d.ns && (sorted[p].info.z & QF_LIVE)) {
        const float4 a = sorted[p].fmin, b = sorted[p].fmax;
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = b.x; v[4] = b.y; v[5] = b.z;
    }
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (uint32_t h = kQueryTile / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 6; ++k) {
                const float o = red[k][threadIdx.x + h];
                // fminf / fmaxf drop a NaN operand: a shape with a non-finite fat box (which no
                // box test accepts) cannot make the tile box NaN and hide the tile's other shapes
                red[k][threadIdx.x] = k < 3 ? fminf(red[k][threadIdx.x], o) : fmaxf(red[k][threadIdx.x], o);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tile_box[2 * blockIdx.x] = make_float4(red[0][0], red[1][0], red[2][0], 0.0f);
        tile_box[2 * blockIdx.x + 1] = make_float4(red[3][0], red[4][0], red[5][0], 0.0f);
    }
}

__device__ __forceinline__ QShape load_qshape(const DevState& d, const float4* geom, const uint4* qinfo, uint32_t j) {
    QShape s;
    s.fmin = d.s_fmin[j];
    s.fmax = d.s_fmax[j];
    s.g0 = geom[3 * j];
    s.g1 = geom[3 * j + 1];
    s.g2 = geom[3 * j + 2];
    s.info = qinfo[j];
    return s;
}

// rays: 2 float4 per ray (origin, max distance; direction, -); hits: 3 float4 per ray (point,
// distance; normal, hit flag bits; body, shape, -, - as bits). Thread i of the grid takes ray
// ray_order[i]; any != 0: stop at the first hit.
__global__ void __launch_bounds__(kQueryTile) k_raycast(DevState d, const QShape* sorted, const float4* tile_box,
                                                        const float4* rays, const uint32_t* ray_order, uint32_t n,
                                                        QueryFilterDev f, uint32_t any, float4* hits) {
    __shared__ QShape tile[kQueryTile];
    const uint32_t i = blockIdx.x * kQueryTile + threadIdx.x;
    const uint32_t ri = i < n ? ray_order[i] : 0u;
    QRay r;
    r.o = mk(0.0f, 0.0f, 0.0f);
    r.d = r.o;
    r.maxd = 0.0f;
    bool active = i < n;
    if (active) {
        const float4 a = rays[2 * ri], b = rays[2 * ri + 1];
        r.o = ld3(a);
        r.maxd = a.w;
        r.d = ld3(b);
    }
    bool found = false;
    float best_t = 0.0f;
    v3 best_n = mk(0.0f, 0.0f, 0.0f);
    uint32_t best_s = 0u, best_owner = 0u;
    const uint32_t ntiles = (d.ns + kQueryTile - 1u) / kQueryTile;
    for (uint32_t tix = 0; tix < ntiles; ++tix) {
        const bool need = active && ray_aabb(r, ld3(tile_box[2 * tix]), ld3(tile_box[2 * tix + 1]));
        if (!__syncthreads_or(need ? 1 : 0)) continue;
        const uint32_t base = tix * kQueryTile;
        if (base + threadIdx.x < d.ns) tile[threadIdx.x] = sorted[base + threadIdx.x];
        __syncthreads();
        const uint32_t m = d.ns - base < kQueryTile ? d.ns - base : kQueryTile;
        for (uint32_t k = 0; need && active && k < m; ++k) {
            const QShape& s = tile[k];
            if (!accepts(s.info, f)) continue;
            if (!ray_aabb(r, ld3(s.fmin), ld3(s.fmax))) continue;
            float t;
            v3 nn;
            if (!shape_ray(r, s, t, nn)) continue;
            const uint32_t slot = __float_as_uint(s.fmin.w);
            if (!found || t < best_t || (t == best_t && slot < best_s)) {
                found = true;
                best_t = t;
                best_n = nn;
                best_s = slot;
                best_owner = s.info.x;
                if (any) active = false;
            }
        }
        __syncthreads();
    }
    if (i >= n) return;
    if (!found) {
        hits[3 * ri] = hits[3 * ri + 1] = hits[3 * ri + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const v3 p = r.o + r.d * best_t;
    hits[3 * ri] = st3(p, best_t);
    hits[3 * ri + 1] = st3(best_n, __uint_as_float(1u));
    hits[3 * ri + 2] = make_float4(__uint_as_float(best_owner), __uint_as_float(best_s), 0.0f, 0.0f);
}

// every hit of one ray, one thread per shape, in arbitrary order (the host sorts them)
__global__ void k_raycast_all(DevState d, const float4* geom, const uint4* qinfo, const float4* ray,
                              QueryFilterDev f, float4* hits, uint32_t cap, uint32_t* count) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.ns) return;
    const QShape s = load_qshape(d, geom, qinfo, j);
    if (!accepts(s.info, f)) return;
    QRay r;
    const float4 a = ray[0], b = ray[1];
    r.o = ld3(a);
    r.maxd = a.w;
    r.d = ld3(b);
    if (!ray_aabb(r, ld3(s.fmin), ld3(s.fmax))) return;
    float t;
    v3 nn;
    if (!shape_ray(r, s, t, nn)) return;
    const uint32_t slot = atomicAdd(count, 1u);
    if (slot >= cap) return;
    const v3 p = r.o + r.d * t;
    hits[3 * slot] = st3(p, t);
    hits[3 * slot + 1] = st3(nn, __uint_as_float(1u));
    hits[3 * slot + 2] = make_float4(__uint_as_float(s.info.x), __uint_as_float(j), 0.0f, 0.0f);
}

// overlap_sphere / overlap_box / overlap_capsule (world.h:1002-1060, query/queries.h:139-258):
// one thread per shape; flags[j] = owner + 1 when shape j overlaps the query, else 0
__device__ __forceinline__ bool shape_overlaps(const QShape& s, const OverlapDev& q) {
    v3 n, p;
    float pen;
    Man m;
    m.count = 0;
    const Sph qs = {ld3(q.c), q.r};
    OBB qb;
    qb.c = ld3(q.c);
    qb.he = ld3(q.he);
    qb.q = ldq(q.q);
    const Cap qc = {ld3(q.a), ld3(q.b), q.r};
    Sph ss;
    ss.c = ld3(s.g0);
    ss.r = s.g0.w;
    OBB sb;
    sb.c = ld3(s.g0);
    sb.he = ld3(s.g1);
    sb.q = ldq(s.g2);
    Cap sc;
    sc.a = ld3(s.g0);
    sc.b = ld3(s.g1);
    sc.r = s.g0.w;
    const uint32_t t = s.info.y;
    if (q.type == SH_SPHERE) {
        if (t == SH_SPHERE) return sphere_sphere(ss, qs, n, pen, p);
        if (t == SH_BOX) return sphere_box_any(qs, sb, false, n, pen, p) != 0;
        return capsule_sphere(sc, qs, n, pen, p);
    }
    if (q.type == SH_BOX) {
        if (t == SH_SPHERE) return sphere_box_any(ss, qb, false, n, pen, p) != 0;
        if (t == SH_BOX) return box_box(sb, qb, m, false);
        return capsule_box(sc, qb, m);
    }
    if (t == SH_SPHERE) return capsule_sphere(qc, ss, n, pen, p);
    if (t == SH_BOX) return capsule_box(qc, sb, m);
    return capsule_capsule(qc, sc, n, pen, p);
}
// the query shape's bounds (world.h:1002-1060: sphere centre +- r, OBB::bounds, capsule end
// points +- r)
__device__ __forceinline__ void overlap_query_bounds(const OverlapDev& q, v3& qmn, v3& qmx) {
    if (q.type == SH_SPHERE) {
        qmn = ld3(q.c) - splat(q.r);
        qmx = ld3(q.c) + splat(q.r);
    } else if (q.type == SH_BOX) {
        OBB qb;
        qb.c = ld3(q.c);
        qb.he = ld3(q.he);
        qb.q = ldq(q.q);
        obb_bounds(qb, qmn, qmx);
    } else {
        qmn = vmin(ld3(q.a), ld3(q.b)) - splat(q.r);
        qmx = vmax(ld3(q.a), ld3(q.b)) + splat(q.r);
    }
}
__global__ void k_overlap(DevState d, const float4* geom, const uint4* qinfo, OverlapDev q, QueryFilterDev f,
                          uint32_t* flags) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.ns) return;
    v3 qmn, qmx;
    overlap_query_bounds(q, qmn, qmx);
    const QShape s = load_qshape(d, geom, qinfo, j);
    uint32_t out = 0u;
    if (accepts(s.info, f)) {
        const v3 fmn = ld3(s.fmin), fmx = ld3(s.fmax);
        // AABB::overlaps (geometry/aabb.h:35-39), inclusive
        const bool box_hit = fmn.x <= qmx.x && fmx.x >= qmn.x && fmn.y <= qmx.y && fmx.y >= qmn.y && fmn.z <= qmx.z &&
                             fmx.z >= qmn.z;
        if (box_hit && shape_overlaps(s, q)) out = s.info.x + 1u;
    }
    flags[j] = out;
}
// one wavefront per 64-slot chunk: the box of the chunk's live fat AABBs (an inverted box when it
// has none), so that a query whose bounds miss it can skip every shape in it
__global__ void __launch_bounds__(256) k_chunk_bounds(DevState d, const uint4* qinfo, float4* chunk_box) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t chunk = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t j = chunk * 64u + lane;
    if (chunk * 64u >= d.ns) return;  // whole wavefronts only; no block barrier
    float v[6] = {3.0e38f, 3.0e38f, 3.0e38f, -3.0e38f, -3.0e38f, -3.0e38f};
    if (j < d.ns && (qinfo[j].z & QF_LIVE)) {
        const float4 a = d.s_fmin[j], b = d.s_fmax[j];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = b.x; v[4] = b.y; v[5] = b.z;
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 6; ++k) {
            const float o = __shfl_xor(v[k], off, 64);
            v[k] = k < 3 ? fminf(v[k], o) : fmaxf(v[k], o);  // NaN-dropping, as in k_tile_bounds
        }
    if (lane == 0) {
        chunk_box[2 * chunk] = make_float4(v[0], v[1], v[2], 0.0f);
        chunk_box[2 * chunk + 1] = make_float4(v[3], v[4], v[5], 0.0f);
    }
}
// A batched overlap query as the kernel takes it: the host's decoded record (type checked on the
// host), or the ABI record read straight from device memory (false for a type other than sphere,
// box or capsule).
__device__ __forceinline__ bool load_overlap(const OverlapDev* qs, uint32_t i, OverlapDev& q) {
    q = qs[i];
    return true;
}
__device__ __forceinline__ bool load_overlap(const pxg_overlap_query* qs, uint32_t i, OverlapDev& q) {
    const pxg_overlap_query r = qs[i];
    q.type = r.type;
    q.r = r.radius;
    q.c = make_float4(r.center[0], r.center[1], r.center[2], 0.0f);
    q.he = make_float4(r.half_extents[0], r.half_extents[1], r.half_extents[2], 0.0f);
    q.q = make_float4(r.orientation[0], r.orientation[1], r.orientation[2], r.orientation[3]);
    q.a = make_float4(r.a[0], r.a[1], r.a[2], 0.0f);
    q.b = make_float4(r.b[0], r.b[1], r.b[2], 0.0f);
    return r.type <= uint32_t(SH_CAPSULE);
}
// Many overlap queries: one wavefront per query scans the shapes in slot order, 64 at a time
// (chunks whose box misses the query's bounds skipped: every live fat AABB of a chunk lies in its
// box, so none of them could pass), and appends the owners of overlapping shapes through a ballot
// (slot order kept); counts[i] is the total, bodies[i * cap ...] holds the first min(cap, total).
// A query record of an unknown shape type gets the count PXG_NO_BODY and no bodies.
template <class Q>
__global__ void __launch_bounds__(256) k_overlap_batch(DevState d, const float4* geom, const uint4* qinfo,
                                                       const float4* chunk_box, const Q* queries, uint32_t n,
                                                       QueryFilterDev f, uint32_t* bodies, uint32_t cap, uint32_t* counts) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t qi = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (qi >= n) return;  // the whole wavefront leaves together; the kernel has no block barrier
    OverlapDev q;
    if (!load_overlap(queries, qi, q)) {
        if (lane == 0) counts[qi] = 0xFFFFFFFFu;
        return;
    }
    v3 qmn, qmx;
    overlap_query_bounds(q, qmn, qmx);
    uint32_t total = 0u;
    const uint32_t chunks = (d.ns + 63u) / 64u;
    // each lane tests one chunk box; the wavefront then visits the chunks of the ballot's set bits
    // in ascending order (one dependent load per 64 chunks instead of one per chunk)
    for (uint32_t cbase = 0; cbase < chunks; cbase += 64u) {
        bool need = false;
        if (cbase + lane < chunks) {
            const v3 cmn = ld3(chunk_box[2 * (cbase + lane)]), cmx = ld3(chunk_box[2 * (cbase + lane) + 1]);
            need = cmn.x <= qmx.x && cmx.x >= qmn.x && cmn.y <= qmx.y && cmx.y >= qmn.y && cmn.z <= qmx.z && cmx.z >= qmn.z;
        }
        unsigned long long visit = __ballot(need ? 1 : 0);
        while (visit) {
            const uint32_t base = (cbase + uint32_t(__ffsll(visit) - 1)) * 64u;
            visit &= visit - 1ull;
            const uint32_t j = base + lane;
            bool hit = false;
            uint32_t owner = 0u;
            if (j < d.ns) {
                const uint4 info = qinfo[j];
                if (accepts(info, f)) {
                    const v3 fmn = ld3(d.s_fmin[j]), fmx = ld3(d.s_fmax[j]);
                    if (fmn.x <= qmx.x && fmx.x >= qmn.x && fmn.y <= qmx.y && fmx.y >= qmn.y && fmn.z <= qmx.z &&
                        fmx.z >= qmn.z) {
                        const QShape s = load_qshape(d, geom, qinfo, j);
                        hit = shape_overlaps(s, q);
                        owner = info.x;
                    }
                }
            }
            const unsigned long long mask = __ballot(hit ? 1 : 0);
            if (hit) {
                const uint32_t k = total + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
                if (k < cap) bodies[size_t(qi) * cap + k] = owner;
            }
            total += uint32_t(__popcll(mask));
        }
    }
    if (lane == 0) counts[qi] = total;
}

// ---- shape casts: sphere, capsule and box (query/sweep.h:68-505, world.h:877-909) ----
constexpr float kSweepTol = 1e-4f;  // sweep.h:68

struct CastCore {  // sweep.h:86-121: core segment (a == b for a sphere) and radius
    v3 a, b;
    float r;
};
__device__ __forceinline__ CastCore cast_core(const ShapeGeom& s, v3 pos, quat rot) {
    CastCore g;
    if (s.type == SH_SPHERE) {
        Sph sp = world_sphere(s, pos, rot);
        g.a = g.b = sp.c;
        g.r = sp.r;
    } else {
        Cap c = world_capsule(s, pos, rot);
        g.a = c.a;
        g.b = c.b;
        g.r = c.r;
    }
    return g;
}
// closest_points_segment_box (collision/distance.h:41-102); returns the squared distance
__device__ float closest_seg_box(v3 p0, v3 p1, const OBB& box, v3& on_seg, v3& on_box) {
    v3 ax[3];
    axes(box, ax);
    const v3 rel0 = p0 - box.c, rel1 = p1 - box.c;
    v3 o, d;
    for (int i = 0; i < 3; ++i) {
        setc(o, i, dot(rel0, ax[i]));
        setc(d, i, dot(rel1, ax[i]) - comp(o, i));
    }
    const v3 h = box.he, nh = -box.he;
    float br[8];
    int count = 0;
    br[count++] = 0.0f;
    br[count++] = 1.0f;
    for (int i = 0; i < 3; ++i) {
        if (fabsf(comp(d, i)) < 1e-12f) continue;
        for (int k = 0; k < 2; ++k) {
            const float plane = k == 0 ? -comp(h, i) : comp(h, i);
            const float sv = (plane - comp(o, i)) / comp(d, i);
            if (sv > 0.0f && sv < 1.0f) br[count++] = sv;
        }
    }
    for (int a = 1; a < count; ++a) {  // sort (equal values are interchangeable)
        const float v = br[a];
        int b = a;
        while (b > 0 && br[b - 1] > v) {
            br[b] = br[b - 1];
            --b;
        }
        br[b] = v;
    }
    auto sq_at = [&](float sv) {
        v3 p = o + d * sv;
        v3 ex = p - mk(gclamp(p.x, nh.x, h.x), gclamp(p.y, nh.y, h.y), gclamp(p.z, nh.z, h.z));
        return dot(ex, ex);
    };
    float best_s = 0.0f;
    float best = sq_at(0.0f);
    for (int k = 0; k + 1 < count; ++k) {
        const float s0 = br[k], s1 = br[k + 1];
        if (s1 - s0 < 1e-9f) continue;
        const v3 mid = o + d * (0.5f * (s0 + s1));
        float a = 0.0f, b = 0.0f;
        for (int i = 0; i < 3; ++i) {
            float c;
            if (comp(mid, i) > comp(h, i)) c = comp(h, i);
            else if (comp(mid, i) < -comp(h, i)) c = -comp(h, i);
            else continue;
            a += comp(d, i) * comp(d, i);
            b += comp(d, i) * (comp(o, i) - c);
        }
        const float sv = a > 1e-12f ? sclamp(-b / a, s0, s1) : s0;
        const float v = sq_at(sv);
        if (v < best) {
            best = v;
            best_s = sv;
        }
    }
    const float end_v = sq_at(1.0f);
    if (end_v < best) {
        best = end_v;
        best_s = 1.0f;
    }
    const v3 local = o + d * best_s;
    const v3 cl = mk(gclamp(local.x, nh.x, h.x), gclamp(local.y, nh.y, h.y), gclamp(local.z, nh.z, h.z));
    on_seg = p0 + (p1 - p0) * best_s;
    on_box = box.c + ax[0] * cl.x + ax[1] * cl.y + ax[2] * cl.z;
    return best;
}
// push_axis_, project_obb_, segment_box_mtv_ (sweep.h:215-230, 304-338)
__device__ __forceinline__ void push_axis(v3* axs, int& count, v3 L) {
    const float len2 = dot(L, L);
    if (len2 < 1e-10f) return;
    axs[count++] = L / sqrtf(len2);
}
__device__ __forceinline__ void project_obb(const OBB& box, const v3 bx[3], v3 L, float& lo, float& hi) {
    const float c = dot(box.c, L);
    const float r = box.he.x * fabsf(dot(bx[0], L)) + box.he.y * fabsf(dot(bx[1], L)) + box.he.z * fabsf(dot(bx[2], L));
    lo = c - r;
    hi = c + r;
}
__device__ void segment_box_mtv(v3 p0, v3 p1, float radius, const OBB& box, v3& normal, float& depth) {
    v3 bx[3];
    axes(box, bx);
    v3 axs[6];
    int count = 0;
    for (int i = 0; i < 3; ++i) push_axis(axs, count, bx[i]);
    for (int i = 0; i < 3; ++i) push_axis(axs, count, cross(p1 - p0, bx[i]));
    depth = 3.402823466e+38f;
    normal = mk(0.0f, 0.0f, 1.0f);
    for (int k = 0; k < count; ++k) {
        const v3 L = axs[k];
        const float s0 = dot(p0, L), s1 = dot(p1, L);
        float b0, b1;
        project_obb(box, bx, L, b0, b1);
        const float lo = gmin(s0, s1), hi = gmax(s0, s1);
        const float overlap = gmin(hi - b0, b1 - lo) + radius;
        if (overlap < depth) {
            depth = overlap;
            normal = (lo + hi) >= (b0 + b1) ? L : -L;
        }
    }
}
// convex_penetration for a sphere or capsule probe (sweep.h:339-384)
__device__ bool convex_penetration(const ShapeGeom& probe, v3 pos, quat rot, const ShapeGeom& target, v3 tpos, quat trot,
                                   v3& normal, float& depth, v3& point) {
    const bool probe_core_vs_box = probe.type == SH_CAPSULE && target.type == SH_BOX;
    const bool box_vs_target_core = probe.type == SH_BOX && target.type == SH_CAPSULE;
    if (probe_core_vs_box || box_vs_target_core) {
        const CastCore core = probe_core_vs_box ? cast_core(probe, pos, rot) : cast_core(target, tpos, trot);
        const OBB box = probe_core_vs_box ? world_obb(target, tpos, trot) : world_obb(probe, pos, rot);
        v3 on_seg, on_box;
        const float dist = sqrtf(closest_seg_box(core.a, core.b, box, on_seg, on_box));
        if (dist >= core.r) return false;
        if (dist > 1e-6f) {
            normal = (on_seg - on_box) / dist;
            depth = core.r - dist;
        } else {
            segment_box_mtv(core.a, core.b, core.r, box, normal, depth);
        }
        point = on_box;
        if (!probe_core_vs_box) normal = -normal;  // the probe is the box: push it away from the core
        return depth > 0.0f;
    }
    Man m;
    if (!generate_contacts(probe, pos, rot, target, tpos, trot, false, m)) return false;
    normal = -m.n;
    depth = 0.0f;
    v3 sum = mk(0.0f, 0.0f, 0.0f);
    for (int k = 0; k < m.count; ++k) {
        depth = gmax(depth, m.p[k].pen);
        sum = sum + m.p[k].pos;
    }
    point = m.count > 0 ? sum / float(m.count) : tpos;
    return depth > 0.0f;
}
// obb_corners_, swept_sat_, sweep_box_box_, polytope_contact_point_ (sweep.h:128-140, 175-258, 282-295)
__device__ void obb_corners(const OBB& box, v3 out[8]) {
    v3 ax[3];
    axes(box, ax);
    int k = 0;
    for (int sx = -1; sx <= 1; sx += 2)
        for (int sy = -1; sy <= 1; sy += 2)
            for (int sz = -1; sz <= 1; sz += 2)
                out[k++] = box.c + ax[0] * (float(sx) * box.he.x) + ax[1] * (float(sy) * box.he.y) +
                           ax[2] * (float(sz) * box.he.z);
}
enum { ADV_HIT = 0, ADV_MISS = 1, ADV_AT_START = 2 };
__device__ int sweep_box_box(const OBB& a, const OBB& b, v3 dir, float maxd, float& toi, v3& axis, v3& mtv_axis,
                             float& mtv_depth) {
    v3 ax[3], bx[3];
    axes(a, ax);
    axes(b, bx);
    v3 axs[15];
    int count = 0;
    for (int i = 0; i < 3; ++i) push_axis(axs, count, ax[i]);
    for (int i = 0; i < 3; ++i) push_axis(axs, count, bx[i]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) push_axis(axs, count, cross(ax[i], bx[j]));
    float t_first = -3.402823466e+38f, t_last = 3.402823466e+38f;
    mtv_depth = 3.402823466e+38f;
    for (int k = 0; k < count; ++k) {
        const v3 L = axs[k];
        float a0, a1, b0, b1;
        project_obb(a, ax, L, a0, a1);
        project_obb(b, bx, L, b0, b1);
        const float v = dot(dir, L);
        const float overlap = gmin(a1 - b0, b1 - a0);
        if (overlap < mtv_depth) {
            mtv_depth = overlap;
            mtv_axis = (a0 + a1) >= (b0 + b1) ? L : -L;
        }
        if (fabsf(v) < 1e-7f) {
            if (overlap <= kSweepTol) return ADV_MISS;
            continue;
        }
        const float enter = v > 0.0f ? (b0 - a1) / v : (b1 - a0) / v;
        const float exit = v > 0.0f ? (b1 - a0) / v : (b0 - a1) / v;
        if (enter > t_first) {
            t_first = enter;
            axis = v > 0.0f ? -L : L;
        }
        t_last = gmin(t_last, exit);
        if (t_first > t_last || t_first > maxd || t_last < 0.0f) return ADV_MISS;
    }
    if (t_first <= kSweepTol) return ADV_AT_START;
    toi = t_first;
    return ADV_HIT;
}
__device__ v3 polytope_contact_point(const OBB& box_at_toi, v3 n, const OBB& target) {
    v3 corners[8];
    obb_corners(box_at_toi, corners);
    float lo = 3.402823466e+38f;
    for (int k = 0; k < 8; ++k) lo = gmin(lo, dot(corners[k], n));
    v3 sum = mk(0.0f, 0.0f, 0.0f);
    int count = 0;
    const float eps = 1e-3f * max4(box_at_toi.he.x, box_at_toi.he.y, box_at_toi.he.z, 1e-3f);
    for (int k = 0; k < 8; ++k) {
        if (dot(corners[k], n) <= lo + eps) {
            sum = sum + corners[k];
            ++count;
        }
    }
    return closest_point_on_obb(sum / float(count > 1 ? count : 1), target);
}
// the gap between the cast shape moved by dir * t and the target (sweep.h:452-483): a sphere or
// capsule core against any target, or a box against a round target's core
__device__ __forceinline__ float cast_gap(bool box_cast, const CastCore& cast, const OBB& cast_box, v3 dir, float t,
                                          const ShapeGeom& target, v3 tpos, quat trot, v3& normal, v3& point) {
    if (box_cast) {
        const CastCore core = cast_core(target, tpos, trot);
        OBB moved = cast_box;
        moved.c = moved.c + dir * t;
        v3 on_seg, on_box;
        const float dist = sqrtf(closest_seg_box(core.a, core.b, moved, on_seg, on_box));
        if (dist > 1e-6f) normal = (on_box - on_seg) / dist;
        point = on_seg + normal * core.r;
        return dist - core.r;
    }
    const v3 a = cast.a + dir * t, b = cast.b + dir * t;
    v3 on_cast, on_target;
    float target_radius = 0.0f;
    if (target.type == SH_BOX) {
        closest_seg_box(a, b, world_obb(target, tpos, trot), on_cast, on_target);
    } else {
        const CastCore core = cast_core(target, tpos, trot);
        seg_seg(a, b, core.a, core.b, on_cast, on_target);
        target_radius = core.r;
    }
    const v3 diff = on_cast - on_target;
    const float dist = len(diff);
    if (dist > 1e-6f) normal = diff / dist;
    point = on_target + normal * target_radius;
    return dist - cast.r - target_radius;
}
// sweep_convex (sweep.h:149-168, 408-416, 422-505)
__device__ bool sweep_convex(const ShapeGeom& shape, v3 pos, quat rot, v3 dir, float maxd, const ShapeGeom& target,
                             v3 tpos, quat trot, float& dist_out, v3& n_out, v3& p_out, bool& started) {
    const bool box_cast = shape.type == SH_BOX;
    const CastCore cast = box_cast ? CastCore{pos, pos, 0.0f} : cast_core(shape, pos, rot);
    const OBB cast_box = world_obb(shape, pos, rot);
    const bool polytopes = box_cast && target.type == SH_BOX;
    v3 n = mk(0.0f, 0.0f, 1.0f), p = mk(0.0f, 0.0f, 0.0f), mtv_axis = mk(0.0f, 0.0f, 1.0f);
    float mtv_depth = 0.0f, toi = 0.0f;
    int result;
    if (polytopes) {
        const OBB tb = world_obb(target, tpos, trot);
        result = sweep_box_box(cast_box, tb, dir, maxd, toi, n, mtv_axis, mtv_depth);
        if (result == ADV_HIT) {
            OBB moved = cast_box;
            moved.c = moved.c + dir * toi;
            p = polytope_contact_point(moved, n, tb);
        } else if (result == ADV_AT_START) {
            p = closest_point_on_obb(cast_box.c, tb);
        }
    } else {
        float t = 0.0f;
        float dd = cast_gap(box_cast, cast, cast_box, dir, t, target, tpos, trot, n, p);
        if (dd <= kSweepTol) {
            result = ADV_AT_START;
        } else {
            result = ADV_HIT;
            for (int iter = 0; iter < 64; ++iter) {
                const float closing = -dot(dir, n);
                if (closing <= 1e-6f) {
                    result = ADV_MISS;
                    break;
                }
                t += dd / closing;
                if (t > maxd) {
                    result = ADV_MISS;
                    break;
                }
                dd = cast_gap(box_cast, cast, cast_box, dir, t, target, tpos, trot, n, p);
                if (dd <= kSweepTol) break;
            }
            toi = t;
        }
    }
    if (result == ADV_MISS) return false;
    if (result == ADV_AT_START) {
        v3 normal = polytopes ? mtv_axis : n, point = p, pen_n, pen_p;
        float depth = polytopes ? mtv_depth : 0.0f, pen_d = 0.0f;
        if (!polytopes && convex_penetration(shape, pos, rot, target, tpos, trot, pen_n, pen_d, pen_p)) {
            normal = pen_n;
            point = pen_p;
            depth = pen_d;
        }
        if (dot(dir, normal) >= -1e-4f) return false;
        dist_out = 0.0f;
        n_out = normal;
        p_out = point;
        started = depth > kSweepTol;
        return true;
    }
    dist_out = toi;
    n_out = n;
    p_out = p;
    started = false;
    return true;
}
__device__ __forceinline__ ShapeGeom decode_shape(const DevState& d, uint32_t i, uint4 info) {
    ShapeGeom s;
    s.type = int(info.y & 15u);
    s.caxis = int((info.y >> 8) & 3u);
    const float4 lc = d.s_lc[i];
    s.lc = mk(lc.x, lc.y, lc.z);
    s.radius = lc.w;
    s.lr = ldq(d.s_lr[i]);
    const float4 dim = d.s_dim[i];
    s.he = mk(dim.x, dim.y, dim.z);
    s.cr = dim.w;
    s.chh = d.s_mat[i].x;
    return s;
}
// one thread per target shape: record (point, distance; normal, started; owner, slot) and the
// nearest (distance, slot) key
__global__ void k_shape_cast(DevState d, const uint4* qinfo, CastDev c, QueryFilterDev f, float4* rec,
                             unsigned long long* best) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.ns) return;
    const float dir_len = len(ld3(c.dir));
    if (!(c.maxd >= 0.0f) || dir_len < 1e-8f) return;
    const v3 ud = ld3(c.dir) / dir_len;
    const v3 pos = ld3(c.pos);
    const quat rot = ldq(c.rot);
    v3 mn0, mx0, mn1, mx1;
    world_bounds(c.shape, pos, rot, mn0, mx0);
    world_bounds(c.shape, pos + ud * c.maxd, rot, mn1, mx1);
    const v3 smn = vmin(mn0, mn1), smx = vmax(mx0, mx1);
    const uint4 qi = qinfo[j];
    if (!accepts(qi, f)) return;
    const v3 fmn = ld3(d.s_fmin[j]), fmx = ld3(d.s_fmax[j]);
    if (!(fmn.x <= smx.x && fmx.x >= smn.x && fmn.y <= smx.y && fmx.y >= smn.y && fmn.z <= smx.z && fmx.z >= smn.z)) return;
    const uint4 info = d.s_info[j];
    const ShapeGeom target = decode_shape(d, j, info);
    const v3 tpos = ld3(d.b_pos[info.x]);
    const quat trot = ldq(d.b_rot[info.x]);
    float dist;
    v3 n, p;
    bool started;
    if (!sweep_convex(c.shape, pos, rot, ud, c.maxd, target, tpos, trot, dist, n, p, started)) return;
    rec[3 * j] = st3(p, dist);
    rec[3 * j + 1] = st3(n, __uint_as_float(started ? 1u : 0u));
    rec[3 * j + 2] = make_float4(__uint_as_float(info.x), __uint_as_float(j), 0.0f, 0.0f);
    atomicMin(best, (static_cast<unsigned long long>(f2o(dist)) << 32) | j);
}
// the nearest record (lower slot on equal distances) as the hit, or an empty hit
__global__ void k_cast_pick(const float4* rec, const unsigned long long* best, float4* hit) {
    const unsigned long long b = *best;
    if (b == ~0ull) {
        hit[0] = hit[1] = hit[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const uint32_t j = uint32_t(b & 0xFFFFFFFFull);
    const float4 r1 = rec[3 * j + 1];
    hit[0] = rec[3 * j];
    hit[1] = make_float4(r1.x, r1.y, r1.z, __uint_as_float(1u));
    const float4 r2 = rec[3 * j + 2];
    hit[2] = make_float4(r2.x, r2.y, r1.w, 0.0f);
}

// Many casts: one wavefront per cast. The lanes share the cast's swept AABB, skip tiles whose box
// misses it, stride over the shapes of the others, sweep each candidate as k_shape_cast does, and
// keep the nearest (distance, slot) key; a wavefront minimum picks the hit, written by the lane
// that holds it (keys are unique: the slot is in the low bits).
constexpr uint32_t kCastWaves = 4;
__global__ void __launch_bounds__(64 * kCastWaves) k_shape_cast_batch(DevState d, const QShape* sorted,
                                                                       const float4* tile_box, const CastDev* casts,
                                                                       uint32_t n, QueryFilterDev f, float4* hits) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t ci = blockIdx.x * kCastWaves + (threadIdx.x >> 6);
    if (ci >= n) return;  // the whole wavefront leaves together; the kernel has no block barrier
    const CastDev c = casts[ci];
    const float dir_len = len(ld3(c.dir));
    unsigned long long best = ~0ull;
    float bd = 0.0f;
    v3 bn = mk(0.0f, 0.0f, 0.0f), bp = bn;
    bool bs = false;
    uint32_t bowner = 0u;
    if (c.maxd >= 0.0f && dir_len >= 1e-8f) {
        const v3 ud = ld3(c.dir) / dir_len;
        const v3 pos = ld3(c.pos);
        const quat rot = ldq(c.rot);
        v3 mn0, mx0, mn1, mx1;
        world_bounds(c.shape, pos, rot, mn0, mx0);
        world_bounds(c.shape, pos + ud * c.maxd, rot, mn1, mx1);
        const v3 smn = vmin(mn0, mn1), smx = vmax(mx0, mx1);
        const uint32_t ntiles = (d.ns + kQueryTile - 1u) / kQueryTile;
        // each lane tests one tile box; the wavefront visits the tiles of the ballot's set bits
        for (uint32_t tbase = 0; tbase < ntiles; tbase += 64u) {
            bool need = false;
            if (tbase + lane < ntiles) {
                const v3 tmn = ld3(tile_box[2 * (tbase + lane)]), tmx = ld3(tile_box[2 * (tbase + lane) + 1]);
                need = tmn.x <= smx.x && tmx.x >= smn.x && tmn.y <= smx.y && tmx.y >= smn.y && tmn.z <= smx.z &&
                       tmx.z >= smn.z;
            }
            unsigned long long visit = __ballot(need ? 1 : 0);
            while (visit) {
                const uint32_t tix = tbase + uint32_t(__ffsll(visit) - 1);
                visit &= visit - 1ull;
                const uint32_t base = tix * kQueryTile;
                const uint32_t m = d.ns - base < kQueryTile ? d.ns - base : kQueryTile;
                for (uint32_t k = lane; k < m; k += 64u) {
                    const QShape& s = sorted[base + k];
                    if (!accepts(s.info, f)) continue;
                    const v3 fmn = ld3(s.fmin), fmx = ld3(s.fmax);
                    if (!(fmn.x <= smx.x && fmx.x >= smn.x && fmn.y <= smx.y && fmx.y >= smn.y && fmn.z <= smx.z &&
                          fmx.z >= smn.z))
                        continue;
                    const uint32_t j = __float_as_uint(s.fmin.w);
                    const uint4 info = d.s_info[j];
                    const ShapeGeom target = decode_shape(d, j, info);
                    float dist;
                    v3 nn, pp;
                    bool started;
                    if (!sweep_convex(c.shape, pos, rot, ud, c.maxd, target, ld3(d.b_pos[info.x]), ldq(d.b_rot[info.x]),
                                      dist, nn, pp, started))
                        continue;
                    const unsigned long long key = (static_cast<unsigned long long>(f2o(dist)) << 32) | j;
                    if (key < best) {
                        best = key;
                        bd = dist;
                        bn = nn;
                        bp = pp;
                        bs = started;
                        bowner = info.x;
                    }
                }
            }
        }
    }
    unsigned long long wmin = best;
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(wmin, off, 64);
        wmin = o < wmin ? o : wmin;
    }
    if (wmin == ~0ull) {
        if (lane == 0) hits[3 * ci] = hits[3 * ci + 1] = hits[3 * ci + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    if (best != wmin) return;
    hits[3 * ci] = st3(bp, bd);
    hits[3 * ci + 1] = st3(bn, __uint_as_float(1u));
    hits[3 * ci + 2] = make_float4(__uint_as_float(bowner), __uint_as_float(uint32_t(best & 0xFFFFFFFFull)),
                                   __uint_as_float(bs ? 1u : 0u), 0.0f);
}

// compute_penetration (world.h:963-995): one thread per target shape; the probe's world bounds
// against the fat AABB (AABB::overlaps, inclusive), the filter, then convex_penetration
__global__ void k_penetration(DevState d, const uint4* qinfo, CastDev c, QueryFilterDev f, float4* rec,
                              uint32_t* count) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.ns) return;
    const uint4 qi = qinfo[j];
    if (!accepts(qi, f)) return;
    const v3 pos = ld3(c.pos);
    const quat rot = ldq(c.rot);
    v3 qmn, qmx;
    world_bounds(c.shape, pos, rot, qmn, qmx);
    const v3 fmn = ld3(d.s_fmin[j]), fmx = ld3(d.s_fmax[j]);
    if (!(fmn.x <= qmx.x && fmx.x >= qmn.x && fmn.y <= qmx.y && fmx.y >= qmn.y && fmn.z <= qmx.z && fmx.z >= qmn.z)) return;
    const uint4 info = d.s_info[j];
    const ShapeGeom target = decode_shape(d, j, info);
    v3 n, p;
    float depth;
    if (!convex_penetration(c.shape, pos, rot, target, ld3(d.b_pos[info.x]), ldq(d.b_rot[info.x]), n, depth, p)) return;
    const uint32_t k = atomicAdd(count, 1u);
    rec[3 * k] = st3(n, depth);
    rec[3 * k + 1] = st3(p, __uint_as_float(j));
    rec[3 * k + 2] = make_float4(__uint_as_float(info.x), 0.0f, 0.0f, 0.0f);
}

// Many compute_penetration queries: one wavefront per posed probe, the shapes in slot order with
// the 64-slot chunk boxes culled as in k_overlap_batch; penetrations are compacted through a
// ballot into out[(i * cap + k) * 10 ...] as pxg_penetration records (normal, depth, point, body,
// shape, 0); counts[i] is the full count.
__global__ void __launch_bounds__(256) k_penetration_batch(DevState d, const uint4* qinfo, const float4* chunk_box,
                                                           const CastDev* probes, uint32_t n, QueryFilterDev f,
                                                           float* out, uint32_t cap, uint32_t* counts) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t qi = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (qi >= n) return;  // the whole wavefront leaves together; the kernel has no block barrier
    const CastDev c = probes[qi];
    const v3 pos = ld3(c.pos);
    const quat rot = ldq(c.rot);
    v3 qmn, qmx;
    world_bounds(c.shape, pos, rot, qmn, qmx);
    const uint32_t chunks = (d.ns + 63u) / 64u;
    uint32_t total = 0u;
    for (uint32_t cbase = 0; cbase < chunks; cbase += 64u) {
        bool need = false;
        if (cbase + lane < chunks) {
            const v3 cmn = ld3(chunk_box[2 * (cbase + lane)]), cmx = ld3(chunk_box[2 * (cbase + lane) + 1]);
            need = cmn.x <= qmx.x && cmx.x >= qmn.x && cmn.y <= qmx.y && cmx.y >= qmn.y && cmn.z <= qmx.z && cmx.z >= qmn.z;
        }
        unsigned long long visit = __ballot(need ? 1 : 0);
        while (visit) {
            const uint32_t j = (cbase + uint32_t(__ffsll(visit) - 1)) * 64u + lane;
            visit &= visit - 1ull;
            bool hit = false;
            v3 n3 = mk(0.0f, 0.0f, 0.0f), p3 = n3;
            float depth = 0.0f;
            uint32_t owner = 0u;
            if (j < d.ns && accepts(qinfo[j], f)) {
                const v3 fmn = ld3(d.s_fmin[j]), fmx = ld3(d.s_fmax[j]);
                if (fmn.x <= qmx.x && fmx.x >= qmn.x && fmn.y <= qmx.y && fmx.y >= qmn.y && fmn.z <= qmx.z &&
                    fmx.z >= qmn.z) {
                    const uint4 info = d.s_info[j];
                    const ShapeGeom target = decode_shape(d, j, info);
                    hit = convex_penetration(c.shape, pos, rot, target, ld3(d.b_pos[info.x]), ldq(d.b_rot[info.x]), n3,
                                             depth, p3);
                    owner = info.x;
                }
            }
            const unsigned long long mask = __ballot(hit ? 1 : 0);
            if (hit) {
                const uint32_t k = total + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
                if (k < cap) {
                    float* o = out + (size_t(qi) * cap + k) * 10u;
                    o[0] = n3.x; o[1] = n3.y; o[2] = n3.z; o[3] = depth;
                    o[4] = p3.x; o[5] = p3.y; o[6] = p3.z;
                    o[7] = __uint_as_float(owner); o[8] = __uint_as_float(j); o[9] = 0.0f;
                }
            }
            total += uint32_t(__popcll(mask));
        }
    }
    if (lane == 0) counts[qi] = total;
}

}  // namespace

void launch_penetration_batch(const DevState& d, const uint4* qinfo, float4* chunk_box, const CastDev* probes, uint32_t n,
                              const QueryFilterDev& f, float* out, uint32_t cap, uint32_t* counts, hipStream_t st) {
    if (!n) return;
    const uint32_t chunks = (d.ns + 63u) / 64u;
    if (chunks) hipLaunchKernelGGL(k_chunk_bounds, dim3((chunks + 3u) / 4u), dim3(256), 0, st, d, qinfo, chunk_box);
    hipLaunchKernelGGL(k_penetration_batch, dim3((n + 3u) / 4u), dim3(256), 0, st, d, qinfo,
                       static_cast<const float4*>(chunk_box), probes, n, f, out, cap, counts);
}

void launch_penetration(const DevState& d, const uint4* qinfo, const CastDev& c, const QueryFilterDev& f, float4* rec,
                        uint32_t* count, hipStream_t st) {
    if (d.ns) hipLaunchKernelGGL(k_penetration, dim3((d.ns + 255u) / 256u), dim3(256), 0, st, d, qinfo, c, f, rec, count);
}

void launch_shape_cast(const DevState& d, const uint4* qinfo, const CastDev& c, const QueryFilterDev& f, float4* rec,
                       unsigned long long* best, float4* hit, hipStream_t st) {
    if (d.ns) hipLaunchKernelGGL(k_shape_cast, dim3((d.ns + 255u) / 256u), dim3(256), 0, st, d, qinfo, c, f, rec, best);
    hipLaunchKernelGGL(k_cast_pick, dim3(1), dim3(1), 0, st, rec, best, hit);
}

void launch_overlap_batch(const DevState& d, const float4* geom, const uint4* qinfo, float4* chunk_box,
                          const OverlapDev* queries, uint32_t n, const QueryFilterDev& f, uint32_t* bodies, uint32_t cap,
                          uint32_t* counts, hipStream_t st) {
    if (!n) return;
    const uint32_t chunks = (d.ns + 63u) / 64u;
    if (chunks) hipLaunchKernelGGL(k_chunk_bounds, dim3((chunks + 3u) / 4u), dim3(256), 0, st, d, qinfo, chunk_box);
    hipLaunchKernelGGL(k_overlap_batch<OverlapDev>, dim3((n + 3u) / 4u), dim3(256), 0, st, d, geom, qinfo,
                       static_cast<const float4*>(chunk_box), queries, n, f, bodies, cap, counts);
}

void launch_overlap_batch_records(const DevState& d, const float4* geom, const uint4* qinfo, float4* chunk_box,
                                  const pxg_overlap_query* queries, uint32_t n, const QueryFilterDev& f, uint32_t* bodies,
                                  uint32_t cap, uint32_t* counts, hipStream_t st) {
    if (!n) return;
    const uint32_t chunks = (d.ns + 63u) / 64u;
    if (chunks) hipLaunchKernelGGL(k_chunk_bounds, dim3((chunks + 3u) / 4u), dim3(256), 0, st, d, qinfo, chunk_box);
    hipLaunchKernelGGL(k_overlap_batch<pxg_overlap_query>, dim3((n + 3u) / 4u), dim3(256), 0, st, d, geom, qinfo,
                       static_cast<const float4*>(chunk_box), queries, n, f, bodies, cap, counts);
}

void launch_overlap(const DevState& d, const float4* geom, const uint4* qinfo, const OverlapDev& q, const QueryFilterDev& f,
                    uint32_t* flags, hipStream_t st) {
    if (d.ns) hipLaunchKernelGGL(k_overlap, dim3((d.ns + 255u) / 256u), dim3(256), 0, st, d, geom, qinfo, q, f, flags);
}

size_t query_sort_bytes(uint32_t cap) {
    size_t bytes = 0;
    uint32_t* none = nullptr;
    if (rocprim::radix_sort_pairs(nullptr, bytes, none, none, none, none, cap, 0u, 30u) != hipSuccess) return 0;
    return bytes;
}
size_t query_shape_record_bytes() { return sizeof(QShape); }
void launch_query_prep(const DevState& d, float4* geom, uint4* qinfo, uint32_t* bounds6, hipStream_t st) {
    if (!d.ns) return;
    hipLaunchKernelGGL(k_query_prep, dim3((d.ns + 255u) / 256u), dim3(256), 0, st, d, geom, qinfo, bounds6);
}
hipError_t launch_shape_cast_batch(const DevState& d, const QueryScratch& q, const CastDev* casts, uint32_t n,
                                   const QueryFilterDev& f, float4* hits, hipStream_t st) {
    if (!n) return hipSuccess;
    if (d.ns) {
        hipLaunchKernelGGL(k_query_keys, dim3((d.ns + 255u) / 256u), dim3(256), 0, st, d, q.info, q.bounds6, q.skey, q.sidx,
                           static_cast<const float4*>(nullptr), 0u, static_cast<uint32_t*>(nullptr),
                           static_cast<uint32_t*>(nullptr));
        size_t bytes = q.sort_bytes;
        hipError_t e = rocprim::radix_sort_pairs(q.sort_tmp, bytes, q.skey, q.skey2, q.sidx, q.sidx2, d.ns, 0u, 30u, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_query_gather, dim3((d.ns + 255u) / 256u), dim3(256), 0, st, d, q.geom, q.info, q.sidx2,
                           static_cast<QShape*>(q.sorted));
        hipLaunchKernelGGL(k_tile_bounds, dim3((d.ns + kQueryTile - 1u) / kQueryTile), dim3(kQueryTile), 0, st, d,
                           static_cast<const QShape*>(q.sorted), q.tile_box);
    }
    hipLaunchKernelGGL(k_shape_cast_batch, dim3((n + kCastWaves - 1u) / kCastWaves), dim3(64 * kCastWaves), 0, st, d,
                       static_cast<const QShape*>(q.sorted), q.tile_box, casts, n, f, hits);
    return hipSuccess;
}

hipError_t launch_raycast(const DevState& d, const QueryScratch& q, const float4* rays, uint32_t n, const QueryFilterDev& f,
                          uint32_t any, float4* hits, hipStream_t st) {
    if (!n) return hipSuccess;
    const uint32_t m = d.ns > n ? d.ns : n;
    hipLaunchKernelGGL(k_query_keys, dim3((m + 255u) / 256u), dim3(256), 0, st, d, q.info, q.bounds6, q.skey, q.sidx, rays,
                       n, q.rkey, q.ridx);
    size_t bytes = q.sort_bytes;
    hipError_t e = hipSuccess;
    if (d.ns) {
        e = rocprim::radix_sort_pairs(q.sort_tmp, bytes, q.skey, q.skey2, q.sidx, q.sidx2, d.ns, 0u, 30u, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_query_gather, dim3((d.ns + 255u) / 256u), dim3(256), 0, st, d, q.geom, q.info, q.sidx2,
                           static_cast<QShape*>(q.sorted));
        hipLaunchKernelGGL(k_tile_bounds, dim3((d.ns + kQueryTile - 1u) / kQueryTile), dim3(kQueryTile), 0, st, d,
                           static_cast<const QShape*>(q.sorted), q.tile_box);
    }
    bytes = q.sort_bytes;
    e = rocprim::radix_sort_pairs(q.sort_tmp, bytes, q.rkey, q.rkey2, q.ridx, q.ridx2, n, 0u, 30u, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_raycast, dim3((n + kQueryTile - 1u) / kQueryTile), dim3(kQueryTile), 0, st, d,
                       static_cast<const QShape*>(q.sorted), q.tile_box, rays, q.ridx2, n, f, any, hits);
    return hipSuccess;
}
void launch_raycast_all(const DevState& d, const float4* geom, const uint4* qinfo, const float4* ray,
                        const QueryFilterDev& f, float4* hits, uint32_t cap, uint32_t* count, hipStream_t st) {
    if (d.ns)
        hipLaunchKernelGGL(k_raycast_all, dim3((d.ns + 255u) / 256u), dim3(256), 0, st, d, geom, qinfo, ray, f, hits, cap,
                           count);
}

}  // namespace pxg
