// XPBD cloth and hair on gfx950: one workgroup (256 threads) advances one cloth or groom by one
// frame (PhysicsWorld::step_cloths_ / step_strands_, world.h:1749-1957), particle positions held
// in LDS for the whole frame. Every stage keeps the reference's results:
//
//  * particle-parallel stages (integration, tethers, collisions, velocities) write only their own
//    particle; constraint batches are particle-disjoint, so a batch projected in parallel equals
//    the reference's serial loop, with a workgroup barrier between batches;
//  * aerodynamic forces are accumulated per particle over its triangles in ascending order (the
//    order the reference's triangle loop adds them in);
//  * anchors, roots and self-collision are applied in the reference's serial order by one lane;
//  * colliders are the world's enabled, non-trigger shapes of live bodies in layer range whose
//    bounds overlap the frame's query box, static bodies' shapes first, each group in shape slot
//    order (the reference visits its two AABB trees; the orders differ only for a particle that
//    meets two colliders of one tree in one substep);
//  * sin, cos, pow and acos are evaluated in double and rounded once.
#include "pxg_soft.h"
#include "pxg_shapes.h"
#include "pxg_joints.h"

using namespace pxd;

namespace pxg {

namespace {

__device__ __forceinline__ float r_sin(float x) { return float(sin(double(x))); }
__device__ __forceinline__ float r_cos(float x) { return float(cos(double(x))); }
__device__ __forceinline__ float r_acos(float x) { return float(acos(double(x))); }
__device__ __forceinline__ float r_pow(float x, float y) { return float(pow(double(x), double(y))); }
__device__ __forceinline__ v3 vmix(v3 x, v3 y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float smix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ v3 safe_normalize(v3 v) {  // util/math.h:46-50
    const float l2 = dot(v, v);
    if (l2 < 1e-6f * 1e-6f) return splat(0.0f);
    return v * (1.0f / sqrtf(l2));
}
__device__ __forceinline__ float wave_maxf(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
// max over the block of a non-negative value (the value only, so any order)
__device__ float block_maxf(float v, float* red) {
    v = wave_maxf(v);
    const uint32_t tid = threadIdx.x;
    if ((tid & 63u) == 0) red[tid >> 6] = v;
    __syncthreads();
    float m = red[0];
    for (uint32_t w = 1; w < kSoftBlock / 64; ++w) m = fmaxf(m, red[w]);
    __syncthreads();
    return m;
}
// compute_bounds (cloth.h:320-327): glm::min / max folded in particle order. Each thread folds a
// contiguous chunk, then one lane folds the chunks in order: the same first-extreme result.
__device__ void fold_bounds(const float4* P, uint32_t n, float4* part_lo, float4* part_hi, v3& lo, v3& hi) {
    const uint32_t tid = threadIdx.x;
    const uint32_t chunk = (n + kSoftBlock - 1) / kSoftBlock;
    v3 a = splat(3.402823466e+38f), b = splat(-3.402823466e+38f);
    for (uint32_t i = tid * chunk; i < n && i < (tid + 1) * chunk; ++i) {
        const v3 p = ld3(P[i]);
        a = vmin(a, p);
        b = vmax(b, p);
    }
    part_lo[tid] = st3(a, 0.0f);
    part_hi[tid] = st3(b, 0.0f);
    __syncthreads();
    if (tid == 0) {
        v3 x = splat(3.402823466e+38f), y = splat(-3.402823466e+38f);
        for (uint32_t t = 0; t < kSoftBlock; ++t) {
            x = vmin(x, ld3(part_lo[t]));
            y = vmax(y, ld3(part_hi[t]));
        }
        part_lo[0] = st3(x, 0.0f);
        part_hi[0] = st3(y, 0.0f);
    }
    __syncthreads();
    lo = ld3(part_lo[0]);
    hi = ld3(part_hi[0]);
    __syncthreads();
}

// The colliders of one cloth or groom, in order, into cols[0..n): per tile of 256 shape slots a
// ballot and a prefix over the 4 waves keep slot order. pass 0: static bodies' shapes; pass 1 the
// rest. `owner` restricts to one body (a groom's collide bodies); kNoBody for any.
struct Gather {
    uint32_t n;
    bool moving, overflow;
};
__device__ void gather_colliders(const DevState& d, uint32_t layer, v3 qlo, v3 qhi, uint32_t owner, int pass, uint32_t* cols,
                                 Gather& g, uint32_t* wave_cnt) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    for (uint32_t base = 0; base < d.ns; base += kSoftBlock) {
        const uint32_t j = base + tid;
        bool take = false, mv = false;
        if (j < d.ns) {
            const uint4 info = d.s_info[j];
            const bool stat = (info.y & SF_STATIC) != 0;
            if ((owner == kNoBody ? (stat == (pass == 0)) : info.x == owner) && (info.y & SF_ENABLED) &&
                !(info.y & SF_TRIGGER) && (d.b_flags[info.x] & BF_ALIVE) &&
                (d.layer_mask[layer & 31u] & (1u << (info.z & 31u)))) {
                const ShapeGeom s = load_shape(d, j, info);
                v3 mn, mx;
                shape_bounds(d, s, ld3(d.b_pos[info.x]), ldq(d.b_rot[info.x]), mn, mx);
                const v3 fmn = ld3(d.s_fmin[j]), fmx = ld3(d.s_fmax[j]);
                const bool in_tree = owner != kNoBody || (fmn.x <= qhi.x && fmx.x >= qlo.x && fmn.y <= qhi.y &&
                                                          fmx.y >= qlo.y && fmn.z <= qhi.z && fmx.z >= qlo.z);
                take = in_tree && mn.x <= qhi.x && mx.x >= qlo.x && mn.y <= qhi.y && mx.y >= qlo.y && mn.z <= qhi.z &&
                       mx.z >= qlo.z;
                if (take && (d.b_flags[info.x] & BF_TYPE) != 0u) {  // world.h:1800-1806: a moving non-static body
                    const v3 v = ld3(d.b_vel[info.x]), w = ld3(d.b_angv[info.x]);
                    mv = dot(v, v) > 1e-6f || dot(w, w) > 1e-6f;
                }
            }
        }
        const unsigned long long m = __ballot(take);
        const uint32_t pre = uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
        if (lane == 0) wave_cnt[wave] = uint32_t(__popcll(m));
        if (__syncthreads_or(mv ? 1 : 0)) g.moving = true;
        uint32_t off = g.n;
        for (uint32_t w = 0; w < wave; ++w) off += wave_cnt[w];
        if (take) {
            if (off + pre < kSoftMaxColliders) cols[off + pre] = j;
            else g.overflow = true;
        }
        uint32_t total = 0;
        for (uint32_t w = 0; w < kSoftBlock / 64; ++w) total += wave_cnt[w];
        g.n += total;
        __syncthreads();
    }
    if (g.n > kSoftMaxColliders) g.n = kSoftMaxColliders;
}

// project_at_pose_ (cloth_collision.h:115-360) with no sweep: p moved `thickness` outside the
// shape; the correction's direction and length
__device__ bool project_particle(const DevState& d, const ShapeGeom& s, v3 pp, quat pr, float thickness, v3& p, v3 prev,
                                 v3& normal, float& depth) {
    if (s.type == SH_SPHERE) {
        const Sph sp = world_sphere(s, pp, pr);
        const v3 dd = p - sp.c;
        const float l = len(dd);
        const float target = sp.r + thickness;
        if (l >= target) return false;
        const v3 n = l > 1e-6f ? dd / l : mk(0.0f, 0.0f, 1.0f);
        depth = target - l;
        p = sp.c + n * target;
        normal = n;
        return true;
    }
    if (s.type == SH_CAPSULE) {
        const Cap cap = world_capsule(s, pp, pr);
        const v3 on = closest_point_on_segment(p, cap.a, cap.b);
        const v3 dd = p - on;
        const float l = len(dd);
        const float target = cap.r + thickness;
        if (l >= target) return false;
        v3 n = dd;
        if (l > 1e-6f) {
            n = dd / l;
        } else {
            v3 t2;
            orthonormal_basis(safe_normalize(cap.b - cap.a), n, t2);
        }
        depth = target - l;
        p = on + n * target;
        normal = n;
        return true;
    }
    if (s.type == SH_BOX) {
        const OBB box = world_obb(s, pp, pr);
        v3 ax[3];
        axes(box, ax);
        const v3 dd = p - box.c;
        const v3 lc = mk(dot(dd, ax[0]), dot(dd, ax[1]), dot(dd, ax[2]));
        const v3 he = box.he;
        v3 lt, ln;
        if (fabsf(lc.x) <= he.x && fabsf(lc.y) <= he.y && fabsf(lc.z) <= he.z) {
            const v3 slack = mk(he.x - fabsf(lc.x), he.y - fabsf(lc.y), he.z - fabsf(lc.z));
            const int a = (slack.x <= slack.y && slack.x <= slack.z) ? 0 : (slack.y <= slack.z ? 1 : 2);
            const float sign = comp(lc, a) >= 0.0f ? 1.0f : -1.0f;
            lt = lc;
            setc(lt, a, sign * (comp(he, a) + thickness));
            ln = splat(0.0f);
            setc(ln, a, sign);
            depth = comp(slack, a) + thickness;
        } else {
            lt = mk(gclamp(lc.x, -he.x, he.x), gclamp(lc.y, -he.y, he.y), gclamp(lc.z, -he.z, he.z));
            const v3 delta = lc - lt;
            const float l = len(delta);
            if (l >= thickness) return false;
            ln = l > 1e-6f ? delta / l : mk(0.0f, 0.0f, 1.0f);
            depth = thickness - l;
            lt = lt + ln * thickness;
        }
        p = box.c + ax[0] * lt.x + ax[1] * lt.y + ax[2] * lt.z;
        normal = normalize(ax[0] * ln.x + ax[1] * ln.y + ax[2] * ln.z);
        return true;
    }
    if (s.type == SH_MESH) {
        const MeshTab t = mesh_tab(d);
        const uint4 meta = t.meta[s.hull];
        const quat rot = qmul(pr, s.lr);
        const quat inv = qinverse(rot);
        const v3 origin = pp + qrot(pr, s.lc);
        const float scale = s.radius > 1e-6f ? s.radius : 1.0f;
        const v3 lp = qrot(inv, p - origin) / scale;
        const float lt = thickness / scale;
        auto corners = [&](uint32_t tri, v3& a, v3& b, v3& c) {
            const uint4 ix = t.tri[meta.y + tri];
            a = ld3(t.v[meta.x + ix.x]);
            b = ld3(t.v[meta.x + ix.y]);
            c = ld3(t.v[meta.x + ix.z]);
        };
        // swept crossing test (cloth_collision.h:243-284)
        const v3 ls = qrot(inv, prev - origin) / scale;
        const v3 seg = lp - ls;
        const float seg_len = len(seg);
        if (seg_len > lt) {
            const v3 dir = seg / seg_len;
            float best_t = seg_len;
            uint32_t crossed_tri = 0;
            bool crossed = false;
            mesh_walk_ray(t, meta, ls, dir, seg_len, [&](uint32_t tri) {
                v3 a, b, c;
                corners(tri, a, b, c);
                float th = 0.0f;
                if ((ray_triangle(ls, dir, seg_len, a, b, c, th) || ray_triangle(ls, dir, seg_len, a, c, b, th)) &&
                    th < best_t) {
                    best_t = th;
                    crossed_tri = tri;
                    crossed = true;
                }
            });
            if (crossed) {
                const v3 hit = ls + dir * best_t;
                v3 face = safe_normalize(ld3(t.n[meta.y + crossed_tri]));
                if (dot(ls - hit, face) < 0.0f) face = -face;
                const v3 fixed = hit + face * lt;
                depth = len(lp - fixed) * scale;
                p = origin + qrot(rot, fixed * scale);
                normal = qrot(rot, face);
                return true;
            }
        }
        // proximity (cloth_collision.h:287-348)
        float best_d2 = lt * lt;
        v3 best_point = splat(0.0f), best_delta = splat(0.0f);
        uint32_t best_tri = 0;
        bool hit = false, any_front = false;
        mesh_walk_box(t, meta, lp - splat(lt), lp + splat(lt), [&](uint32_t tri) {
            v3 a, b, c;
            corners(tri, a, b, c);
            const v3 cl = closest_on_triangle(lp, a, b, c);
            const v3 delta = lp - cl;
            const float d2 = dot(delta, delta);
            if (d2 > lt * lt) return;
            if (dot(delta, ld3(t.n[meta.y + tri])) >= 0.0f) any_front = true;
            if (d2 > best_d2) return;
            best_d2 = d2;
            best_point = cl;
            best_delta = delta;
            best_tri = tri;
            hit = true;
        });
        if (!hit) return false;
        v3 bn;
        if (!any_front) {
            bn = safe_normalize(ld3(t.n[meta.y + best_tri]));
        } else if (best_d2 > 1e-6f * 1e-6f) {
            bn = best_delta / sqrtf(best_d2);
        } else {
            const v3 face = safe_normalize(ld3(t.n[meta.y + best_tri]));
            bn = dot(best_delta, face) < 0.0f ? -face : face;
        }
        const v3 fixed = best_point + bn * lt;
        depth = len(lp - fixed) * scale;
        p = origin + qrot(rot, fixed * scale);
        normal = qrot(rot, bn);
        return true;
    }
    if (s.type == SH_HULL) {
        const HullTab ht = hull_tab(d);
        Sph probe;
        probe.c = p;
        probe.r = thickness;
        v3 point;
        float pen = 0.0f;
        if (!hull_sphere(ht, posed_hull(ht, s.hull, pp, pr, s), probe, normal, pen, point)) return false;
        p = point + normal * thickness;
        depth = pen;
        return true;
    }
    return false;
}

// detail::project_distance (cloth_solver.h:103-121) on LDS positions
__device__ __forceinline__ void project_distance(float4* P, uint4* con, uint32_t k, float alpha) {
    const uint4 c = con[k];
    const float4 pa = P[c.x], pb = P[c.y];
    const float w = pa.w + pb.w;
    if (w <= 0.0f) return;
    const v3 dd = ld3(pa) - ld3(pb);
    const float l = len(dd);
    if (l < 1e-6f) return;
    const v3 n = dd / l;
    const float cc = l - __uint_as_float(c.z);
    const float lam = __uint_as_float(c.w);
    const float dl = (-cc - alpha * lam) / (w + alpha);
    con[k].w = __float_as_uint(lam + dl);
    const v3 corr = dl * n;
    P[c.x] = st3(ld3(pa) + pa.w * corr, pa.w);
    P[c.y] = st3(ld3(pb) - pb.w * corr, pb.w);
}

__device__ __forceinline__ uint32_t cloth_hash(int32_t x, int32_t y, int32_t z, uint32_t buckets) {  // cloth_solver.h:88-93
    const uint32_t h = uint32_t(x) * 73856093u ^ uint32_t(y) * 19349663u ^ uint32_t(z) * 83492791u;
    return h & (buckets - 1u);
}

// step_cloths_ (world.h:1749-1847) + solve_cloth (cloth_solver.h:126-411) for cloth blockIdx.x
__global__ __launch_bounds__(kSoftBlock) void k_cloth(DevState d, SoftDev s, float dt, uint32_t frame_substeps,
                                                      uint32_t world_iters, v3 gravity, float margin) {
    extern __shared__ float4 P[];  // positions, w = inv_mass
    __shared__ float4 part_lo[kSoftBlock], part_hi[kSoftBlock];
    __shared__ float red[kSoftBlock / 64];
    __shared__ uint32_t wave_cnt[kSoftBlock / 64];
    const uint32_t ci = blockIdx.x, tid = threadIdx.x;
    const ClothRec r = s.cl[ci];
    if (!r.alive || !r.enabled || r.np == 0) return;
    SoftState stt = s.cl_state[ci];
    float4* pos = s.cl_pos + r.p0;
    float4* prv = s.cl_prev + r.p0;
    float4* vel = s.cl_vel + r.p0;
    float4* anc = s.cl_anchor + 3 * size_t(r.a0);
    uint4* con = s.cl_con + r.c0;
    for (uint32_t i = tid; i < r.np; i += kSoftBlock) P[i] = pos[i];
    // anchors follow their bodies (world.h:1764-1772)
    bool moved = false;
    for (uint32_t a = tid; a < r.n_anchor; a += kSoftBlock) {
        const float4 w = anc[3 * a], lc = anc[3 * a + 2];
        const uint32_t body = s.cl_anchor_body[r.a0 + a];
        const v3 target = body != kNoBody ? ld3(d.b_pos[body]) + qrot(ldq(d.b_rot[body]), ld3(lc)) : ld3(w);
        const v3 delta = target - ld3(w);
        if (dot(delta, delta) > 1e-6f) moved = true;
        anc[3 * a] = st3(target, w.w);
    }
    const bool anchor_moved = __syncthreads_or(moved ? 1 : 0) != 0;
    // candidate colliders (world.h:1774-1810)
    float m2 = 0.0f;
    for (uint32_t i = tid; i < r.np; i += kSoftBlock) {
        const v3 v = ld3(vel[i]);
        m2 = fmaxf(m2, dot(v, v));
    }
    const float travel = sqrtf(block_maxf(m2, red)) * dt;
    const float grow = (r.thickness + travel) + margin;
    const v3 qlo = mk(stt.bmin[0], stt.bmin[1], stt.bmin[2]) - splat(grow);
    const v3 qhi = mk(stt.bmax[0], stt.bmax[1], stt.bmax[2]) + splat(grow);
    uint32_t* cols = s.cl_cols + size_t(ci) * kSoftMaxColliders;
    Gather g{0u, false, false};
    gather_colliders(d, r.layer, qlo, qhi, kNoBody, 0, cols, g, wave_cnt);
    gather_colliders(d, r.layer, qlo, qhi, kNoBody, 1, cols, g, wave_cnt);
    const bool overflow = __syncthreads_or(g.overflow ? 1 : 0) != 0;
    __syncthreads();
    if (anchor_moved || g.moving) {  // the sleep gate (world.h:1812-1823)
        stt.awake = 1u;
        stt.sleep_timer = 0.0f;
    }
    if (!stt.awake) {
        for (uint32_t a = tid; a < r.n_anchor; a += kSoftBlock) {
            const float4 w = anc[3 * a], pw = anc[3 * a + 1];
            anc[3 * a + 1] = make_float4(w.x, w.y, w.z, pw.w);
        }
        if (tid == 0) {
            stt.error = overflow ? 1u : 0u;
            s.cl_state[ci] = stt;
        }
        return;
    }
    // ---- solve_cloth ----
    const uint32_t steps = max(1u, r.substeps ? r.substeps : frame_substeps);
    const uint32_t iters = max(1u, r.iterations ? r.iterations : world_iters);
    const float hs = dt / float(steps);
    const float inv_hs = 1.0f / hs;
    const bool self_col = r.self_collision && r.self_distance > 1e-6f;
    const float s_alpha = r.stretch_compliance / (hs * hs);
    const float b_alpha = r.bend_compliance / (hs * hs);
    v3 wind = mk(r.wind[0], r.wind[1], r.wind[2]);
    if (r.wind_turbulence > 0.0f) {
        const float t = stt.time;
        wind = wind + r.wind_turbulence * mk(r_sin(t * 1.7f) * r_cos(t * 0.9f), r_sin(t * 1.1f + 1.3f) * r_cos(t * 1.9f),
                                             r_sin(t * 0.7f + 2.7f) * r_cos(t * 1.3f));
    }
    const v3 base = gravity * r.gravity_scale + mk(r.ext[0], r.ext[1], r.ext[2]);
    const bool aero = (r.air_drag > 0.0f || r.air_lift > 0.0f) && r.n_tri > 0;
    const float maxv2 = r.max_velocity * r.max_velocity;
    const float damp = 1.0f / (1.0f + r.damping * hs);
    const uint32_t n_col = g.n;
    float time = stt.time;
    for (uint32_t step = 0; step < steps; ++step) {
        if (aero) {  // 1. aerodynamics
            for (uint32_t t = tid; t < r.n_tri; t += kSoftBlock) {
                const uint4 tr = s.cl_tri[r.tr0 + t];
                const v3 p0 = ld3(P[tr.x]);
                const v3 e1 = ld3(P[tr.y]) - p0, e2 = ld3(P[tr.z]) - p0;
                const v3 cr = cross(e1, e2);
                const float twice = len(cr);
                v3 share = splat(0.0f);
                const bool live = !(twice < 1e-6f);  // the reference skips only twice < epsilon
                if (live) {
                    const v3 fn = cr / twice;
                    const v3 v = (ld3(vel[tr.x]) + ld3(vel[tr.y]) + ld3(vel[tr.z])) * (1.0f / 3.0f) - wind;
                    const float vn = dot(v, fn);
                    const v3 vt = v - vn * fn;
                    const v3 force = -0.5f * twice * fabsf(vn) * (r.air_drag * vn * fn + r.air_lift * vt);
                    share = force * (1.0f / 3.0f);
                }
                s.cl_share[r.tr0 + t] = st3(share, live ? 1.0f : 0.0f);
            }
            __syncthreads();
        }
        for (uint32_t i = tid; i < r.np; i += kSoftBlock) {  // 2. integrate
            const float4 pi = P[i];
            prv[i] = pi;
            if (pi.w == 0.0f) continue;
            v3 acc = splat(0.0f);
            if (aero) {
                const uint32_t b = s.cl_ptri_off[r.pt0 + i], e = s.cl_ptri_off[r.pt0 + i + 1];
                for (uint32_t k = b; k < e; ++k) {
                    const float4 sh = s.cl_share[r.tr0 + s.cl_ptri[k]];
                    if (sh.w != 0.0f) acc = acc + ld3(sh);
                }
            }
            v3 v = ld3(vel[i]);
            v = v + (base + acc * pi.w) * hs;
            v = v * damp;
            vel[i] = st3(v, 0.0f);
            P[i] = st3(ld3(pi) + v * hs, pi.w);
        }
        __syncthreads();
        if (tid == 0) {  // 3. anchors, in order
            const float at = float(step + 1) / float(steps);
            for (uint32_t a = 0; a < r.n_anchor; ++a) {
                const float4 w = anc[3 * a], pw = anc[3 * a + 1];
                const uint32_t pi = __float_as_uint(w.w);
                if (pi >= r.np) continue;
                P[pi] = st3(vmix(ld3(pw), ld3(w), at), P[pi].w);
            }
        }
        for (uint32_t k = tid; k < r.n_stretch + r.n_bend; k += kSoftBlock) con[k].w = 0u;  // 4. constraints
        __syncthreads();
        for (uint32_t it = 0; it < iters; ++it) {
            for (uint32_t bi = 0; bi < r.n_sb + r.n_bb; ++bi) {
                const uint2 b = s.cl_batch[r.b0 + bi];
                const float alpha = bi < r.n_sb ? s_alpha : b_alpha;
                for (uint32_t k = b.x + tid; k < b.y; k += kSoftBlock) project_distance(P, con, k, alpha);
                __syncthreads();
            }
        }
        const uint32_t t_first = r.tether_serial ? (tid == 0 ? 0u : r.n_tether) : tid;
        const uint32_t t_stride = r.tether_serial ? 1u : kSoftBlock;
        for (uint32_t t = t_first; t < r.n_tether; t += t_stride) {  // 5. tethers
            const uint4 te = s.cl_tether[r.t0 + t];
            const float4 pp = P[te.x];
            if (pp.w == 0.0f) continue;
            const v3 anchor = ld3(P[te.y]);
            const v3 dd = ld3(pp) - anchor;
            const float l = len(dd);
            const float ml = __uint_as_float(te.z);
            if (l <= ml || l < 1e-6f) continue;
            P[te.x] = st3(anchor + dd * (ml / l), pp.w);
        }
        __syncthreads();
        if (self_col && tid == 0) {  // 6. self-collision, serial (cloth_solver.h:289-361)
            const float dmin = r.self_distance;
            const float inv_cell = 1.0f / dmin;
            const uint32_t nb = r.buckets;
            uint32_t* cell_of = s.cl_bins + r.bins0;
            uint32_t* counts = cell_of + r.np;
            uint32_t* sorted = counts + nb + 1;
            for (uint32_t b = 0; b <= nb; ++b) counts[b] = 0u;
            for (uint32_t i = 0; i < r.np; ++i) {
                const v3 p = ld3(P[i]);
                const uint32_t c = cloth_hash(int32_t(floorf(p.x * inv_cell)), int32_t(floorf(p.y * inv_cell)),
                                              int32_t(floorf(p.z * inv_cell)), nb);
                cell_of[i] = c;
                ++counts[c];
            }
            uint32_t run = 0;
            for (uint32_t b = 0; b < nb; ++b) {
                const uint32_t k = counts[b];
                counts[b] = run;
                run += k;
            }
            counts[nb] = run;
            for (uint32_t i = 0; i < r.np; ++i) sorted[counts[cell_of[i]]++] = i;
            for (uint32_t i = 0; i < r.np; ++i) {
                float4 pi = P[i];
                if (pi.w == 0.0f) continue;
                const v3 bc = ld3(pi) * inv_cell;
                const int32_t cx = int32_t(floorf(bc.x)), cy = int32_t(floorf(bc.y)), cz = int32_t(floorf(bc.z));
                for (int32_t dz = -1; dz <= 1; ++dz)
                    for (int32_t dy = -1; dy <= 1; ++dy)
                        for (int32_t dx = -1; dx <= 1; ++dx) {
                            const uint32_t b = cloth_hash(cx + dx, cy + dy, cz + dz, nb);
                            const uint32_t first = b == 0 ? 0u : counts[b - 1], last = counts[b];
                            for (uint32_t q = first; q < last; ++q) {
                                const uint32_t j = sorted[q];
                                if (j <= i) continue;
                                if (r.columns > 0) {
                                    const int32_t xi = int32_t(i % r.columns), yi = int32_t(i / r.columns);
                                    const int32_t xj = int32_t(j % r.columns), yj = int32_t(j / r.columns);
                                    if (abs(xi - xj) <= 1 && abs(yi - yj) <= 1) continue;
                                }
                                const float4 pj = P[j];
                                const float w = pi.w + pj.w;
                                if (w <= 0.0f) continue;
                                const v3 dd = ld3(pi) - ld3(pj);
                                const float l2 = dot(dd, dd);
                                if (l2 >= dmin * dmin || l2 < 1e-6f) continue;
                                const float l = sqrtf(l2);
                                const v3 push = (dd / l) * ((dmin - l) / w);
                                pi = st3(ld3(pi) + push * pi.w, pi.w);
                                P[j] = st3(ld3(pj) - push * pj.w, pj.w);
                            }
                        }
                P[i] = pi;
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < r.np; i += kSoftBlock) {  // 7. collisions, 8. velocities and friction
            float4 pi = P[i];
            const v3 pv = ld3(prv[i]);
            v3 p = ld3(pi);
            v3 nsum = splat(0.0f);
            float depth = 0.0f;
            if (pi.w != 0.0f) {
                for (uint32_t c = 0; c < n_col; ++c) {
                    const uint32_t slot = cols[c];
                    const uint4 info = d.s_info[slot];
                    const ShapeGeom sh = load_shape(d, slot, info);
                    const v3 bp = ld3(d.b_pos[info.x]);
                    const quat bq = ldq(d.b_rot[info.x]);
                    v3 mn, mx;
                    shape_bounds(d, sh, bp, bq, mn, mx);
                    const float m = r.thickness;
                    if (p.x < mn.x - m || p.x > mx.x + m || p.y < mn.y - m || p.y > mx.y + m || p.z < mn.z - m ||
                        p.z > mx.z + m)
                        continue;
                    v3 nrm = splat(0.0f);
                    float dep = 0.0f;
                    if (!project_particle(d, sh, bp, bq, r.thickness, p, pv, nrm, dep)) continue;
                    nsum = nsum + nrm;
                    depth += dep;
                }
            }
            v3 v = (p - pv) * inv_hs;
            if (depth > 0.0f) {
                const float nl = len(nsum);
                if (nl > 1e-6f) {
                    const v3 nrm = nsum / nl;
                    const float vn = dot(v, nrm);
                    const v3 vt = v - vn * nrm;
                    if (vn < 0.0f) v = v - vn * nrm;
                    const float budget = r.friction * depth * inv_hs;
                    const float vt_len = len(vt);
                    if (vt_len > 1e-6f) {
                        const float q = budget / vt_len;
                        v = v - vt * ((q < 1.0f) ? q : 1.0f);
                    }
                }
            }
            const float v2 = dot(v, v);
            if (v2 > maxv2) v = v * (r.max_velocity / sqrtf(v2));
            vel[i] = st3(v, 0.0f);
            P[i] = st3(p, pi.w);
        }
        time += hs;
        __syncthreads();
    }
    for (uint32_t a = tid; a < r.n_anchor; a += kSoftBlock) {
        const float4 w = anc[3 * a], pw = anc[3 * a + 1];
        anc[3 * a + 1] = make_float4(w.x, w.y, w.z, pw.w);
    }
    for (uint32_t i = tid; i < r.np; i += kSoftBlock) pos[i] = P[i];
    v3 lo, hi;
    fold_bounds(P, r.np, part_lo, part_hi, lo, hi);
    // sleep bookkeeping (world.h:1832-1845)
    float s2 = 0.0f;
    for (uint32_t i = tid; i < r.np; i += kSoftBlock) {
        const v3 v = ld3(vel[i]);
        s2 = fmaxf(s2, dot(v, v));
    }
    const float speed = sqrtf(block_maxf(s2, red));
    bool sleep_now = false;
    if (speed < r.sleep_threshold) {
        stt.sleep_timer += dt;
        if (stt.sleep_timer >= r.sleep_time) {
            stt.awake = 0u;
            sleep_now = true;
        }
    } else {
        stt.sleep_timer = 0.0f;
    }
    if (sleep_now)
        for (uint32_t i = tid; i < r.np; i += kSoftBlock) vel[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (tid == 0) {
        stt.time = time;
        stt.bmin[0] = lo.x; stt.bmin[1] = lo.y; stt.bmin[2] = lo.z;
        stt.bmax[0] = hi.x; stt.bmax[1] = hi.y; stt.bmax[2] = hi.z;
        stt.error = overflow ? 1u : 0u;
        s.cl_state[ci] = stt;
    }
}

// glm::angleAxis and glm::slerp (0.9.9), hair's rotation_between (strand_solver.h:70-85)
__device__ quat angle_axis(float angle, v3 axis) {
    const float sn = r_sin(angle * 0.5f);
    quat q;
    q.w = r_cos(angle * 0.5f);
    q.x = axis.x * sn; q.y = axis.y * sn; q.z = axis.z * sn;
    return q;
}
__device__ quat slerp(quat x, quat y, float a) {
    quat z = y;
    float ct = (x.x * y.x + x.y * y.y) + (x.z * y.z + x.w * y.w);
    if (ct < 0.0f) {
        z.w = -y.w; z.x = -y.x; z.y = -y.y; z.z = -y.z;
        ct = -ct;
    }
    quat o;
    if (ct > 1.0f - 1.1920928955078125e-7f) {
        o.w = smix(x.w, z.w, a); o.x = smix(x.x, z.x, a); o.y = smix(x.y, z.y, a); o.z = smix(x.z, z.z, a);
        return o;
    }
    const float angle = r_acos(ct);
    const float s1 = r_sin((1.0f - a) * angle), s2 = r_sin(a * angle), sd = r_sin(angle);
    o.w = (x.w * s1 + z.w * s2) / sd;
    o.x = (x.x * s1 + z.x * s2) / sd;
    o.y = (x.y * s1 + z.y * s2) / sd;
    o.z = (x.z * s1 + z.z * s2) / sd;
    return o;
}
__device__ quat rotation_between(v3 from, v3 to) {
    const float dd = gclamp(dot(from, to), -1.0f, 1.0f);
    quat id;
    id.w = 1.0f; id.x = id.y = id.z = 0.0f;
    if (dd > 0.999999f) return id;
    if (dd < -0.999999f) {
        v3 axis = cross(mk(0.0f, 0.0f, 1.0f), from);
        if (dot(axis, axis) < 1e-8f) axis = cross(mk(1.0f, 0.0f, 0.0f), from);
        return angle_axis(3.14159265358979323846f, normalize(axis));
    }
    const v3 axis = cross(from, to);
    const float l = len(axis);
    if (l < 1e-8f) return id;
    return angle_axis(r_acos(dd), axis / l);
}

// step_strands_ (world.h:1861-1957) + solve_strands (strand_solver.h:121-317) for set blockIdx.x
__global__ __launch_bounds__(kSoftBlock) void k_strands(DevState d, SoftDev s, float dt, uint32_t frame_substeps,
                                                        uint32_t world_iters, v3 gravity, float margin) {
    extern __shared__ float4 P[];
    __shared__ float red[kSoftBlock / 64];
    __shared__ uint32_t wave_cnt[kSoftBlock / 64];
    __shared__ float4 part_lo[kSoftBlock], part_hi[kSoftBlock];
    const uint32_t si = blockIdx.x, tid = threadIdx.x;
    const StrandRec r = s.st[si];
    if (!r.alive || !r.enabled || r.np == 0 || r.n_strands == 0) return;
    SoftState stt = s.st_state[si];
    float4* pos = s.st_pos + r.p0;
    float4* prv = s.st_prev + r.p0;
    float4* vel = s.st_vel + r.p0;
    uint4* con = s.st_con + r.c0;
    float4* strand = s.st_strand + 4 * size_t(r.s0);
    for (uint32_t i = tid; i < r.np; i += kSoftBlock) P[i] = pos[i];
    bool moved = false;
    for (uint32_t k = tid; k < r.n_strands; k += kSoftBlock) {
        const v3 dd = ld3(strand[4 * k]) - ld3(strand[4 * k + 2]);
        if (dot(dd, dd) > 1e-6f) moved = true;
    }
    const bool root_moved = __syncthreads_or(moved ? 1 : 0) != 0;
    float m2 = 0.0f;
    for (uint32_t i = tid; i < r.np; i += kSoftBlock) {
        const v3 v = ld3(vel[i]);
        m2 = fmaxf(m2, dot(v, v));
    }
    const float travel = sqrtf(block_maxf(m2, red)) * dt;
    const float grow = (r.thickness + travel) + margin;
    const v3 qlo = mk(stt.bmin[0], stt.bmin[1], stt.bmin[2]) - splat(grow);
    const v3 qhi = mk(stt.bmax[0], stt.bmax[1], stt.bmax[2]) + splat(grow);
    uint32_t* cols = s.st_cols + size_t(si) * kSoftMaxColliders;
    Gather g{0u, false, false};
    if (r.n_cb == 0) {
        gather_colliders(d, r.layer, qlo, qhi, kNoBody, 0, cols, g, wave_cnt);
        gather_colliders(d, r.layer, qlo, qhi, kNoBody, 1, cols, g, wave_cnt);
    } else {
        for (uint32_t k = 0; k < r.n_cb; ++k) {
            const uint32_t body = s.st_collide[r.cb0 + k];
            if (body != kNoBody) gather_colliders(d, r.layer, qlo, qhi, body, 0, cols, g, wave_cnt);
        }
    }
    const bool overflow = __syncthreads_or(g.overflow ? 1 : 0) != 0;
    __syncthreads();
    if (root_moved || g.moving) {
        stt.awake = 1u;
        stt.sleep_timer = 0.0f;
    }
    if (!stt.awake) {
        for (uint32_t k = tid; k < r.n_strands; k += kSoftBlock) {
            const float4 a = strand[4 * k], q = strand[4 * k + 1], b = strand[4 * k + 2];
            strand[4 * k + 2] = make_float4(a.x, a.y, a.z, b.w);
            strand[4 * k + 3] = q;
        }
        if (tid == 0) {
            stt.error = overflow ? 1u : 0u;
            s.st_state[si] = stt;
        }
        return;
    }
    const uint32_t steps = max(1u, r.substeps ? r.substeps : frame_substeps);
    const uint32_t iters = max(1u, r.iterations ? r.iterations : world_iters);
    const float hs = dt / float(steps);
    const float inv_hs = 1.0f / hs;
    const float alpha = r.stretch_compliance * inv_hs * inv_hs;
    const v3 accel = gravity * r.gravity_scale + mk(r.ext[0], r.ext[1], r.ext[2]);
    const float maxv2 = r.max_velocity * r.max_velocity;
    const float frac = hs / (1.0f / 60.0f);
    const float g_sub = 1.0f - r_pow(1.0f - gclamp(r.global_shape, 0.0f, 1.0f), frac);
    const float l_sub = 1.0f - r_pow(1.0f - gclamp(r.local_shape, 0.0f, 1.0f), frac);
    const v3 wind0 = mk(r.wind[0], r.wind[1], r.wind[2]);
    const float wind_len = len(wind0);
    const uint32_t n_col = g.n;
    float time = stt.time;
    for (uint32_t step = 0; step < steps; ++step) {
        const float t = float(step + 1) / float(steps);
        for (uint32_t i = tid; i < r.np; i += kSoftBlock) {  // 1 + 2: aerodynamics and integration
            const float4 pi = P[i];
            if (pi.w <= 0.0f) continue;
            v3 v = ld3(vel[i]);
            v3 a = accel;
            if (r.air_drag > 0.0f) {
                v3 wind = wind0;
                if (r.wind_turbulence > 0.0f) {
                    const float ph = ((time * 2.3f + pi.x * 1.7f) + pi.y * 2.1f) + pi.z * 1.3f;
                    wind = wind + r.wind_turbulence * wind_len * mk(r_sin(ph), r_sin(ph * 1.37f + 2.1f), r_sin(ph * 0.81f + 4.2f));
                }
                a = a + (wind - v) * r.air_drag;
            }
            v = v + a * hs;
            v = v * (1.0f / (1.0f + r.damping * hs));
            const float v2 = dot(v, v);
            if (v2 > maxv2) v = v * (r.max_velocity / sqrtf(v2));
            vel[i] = st3(v, 0.0f);
            prv[i] = pi;
            P[i] = st3(ld3(pi) + v * hs, pi.w);
        }
        __syncthreads();
        for (uint32_t k = tid; k < r.n_strands; k += kSoftBlock) {  // 3: roots
            const float4 root = strand[4 * k], proot = strand[4 * k + 2];
            const uint32_t first = __float_as_uint(root.w), count = __float_as_uint(proot.w);
            if (count == 0) continue;
            prv[first] = P[first];
            P[first] = st3(vmix(ld3(proot), ld3(root), t), P[first].w);
            vel[first] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        for (uint32_t k = tid; k < r.n_stretch; k += kSoftBlock) con[k].w = 0u;  // 4: stretch
        __syncthreads();
        for (uint32_t it = 0; it < iters; ++it)
            for (uint32_t bi = 0; bi < r.n_b; ++bi) {
                const uint2 b = s.st_batch[r.b0 + bi];
                for (uint32_t k = b.x + tid; k < b.y; k += kSoftBlock) project_distance(P, con, k, alpha);
                __syncthreads();
            }
        if (r.global_shape > 0.0f) {  // 5: global shape, one lane per strand
            for (uint32_t k = tid; k < r.n_strands; k += kSoftBlock) {
                const float4 root4 = strand[4 * k], proot4 = strand[4 * k + 2];
                const uint32_t first = __float_as_uint(root4.w), count = __float_as_uint(proot4.w);
                if (count < 2) continue;
                const quat q = slerp(ldq(strand[4 * k + 3]), ldq(strand[4 * k + 1]), t);
                const v3 root = vmix(ld3(proot4), ld3(root4), t);
                const float inv_last = 1.0f / float(count - 1);
                for (uint32_t j = 1; j < count; ++j) {
                    const uint32_t i = first + j;
                    const float4 pi = P[i];
                    if (pi.w <= 0.0f) continue;
                    const float along = float(j) * inv_last;
                    const float weight = g_sub * r_pow(1.0f - along, r.global_shape_falloff);
                    if (weight <= 0.0f) continue;
                    const v3 target = root + qrot(q, ld3(s.st_rest_local[r.p0 + i]));
                    P[i] = st3(vmix(ld3(pi), target, gmin(weight, 1.0f)), pi.w);
                }
            }
            __syncthreads();
        }
        if (r.local_shape > 0.0f) {  // 6: local shape, one lane per strand in order
            for (uint32_t k = tid; k < r.n_strands; k += kSoftBlock) {
                const uint32_t first = __float_as_uint(strand[4 * k].w), count = __float_as_uint(strand[4 * k + 2].w);
                if (count < 3) continue;
                for (uint32_t j = 1; j + 1 < count; ++j) {
                    const uint32_t prev = first + j - 1, cur = first + j, next = first + j + 1;
                    const v3 rest_prev = ld3(s.st_rest_offset[r.p0 + cur]);
                    const v3 live_prev = ld3(P[cur]) - ld3(P[prev]);
                    const float rl = len(rest_prev), ll = len(live_prev);
                    if (rl <= 1e-8f || ll <= 1e-8f) continue;
                    const quat turn = rotation_between(rest_prev / rl, live_prev / ll);
                    const v3 target = ld3(P[cur]) + qrot(turn, ld3(s.st_rest_offset[r.p0 + next]));
                    const v3 delta = (target - ld3(P[next])) * l_sub;
                    const float4 a = P[cur], b = P[next];
                    const float w = a.w + b.w;
                    if (w <= 0.0f) continue;
                    P[next] = st3(ld3(b) + delta * (b.w / w), b.w);
                    P[cur] = st3(ld3(a) - delta * (a.w / w), a.w);
                }
            }
            __syncthreads();
        }
        for (uint32_t i = tid; i < r.np; i += kSoftBlock) {  // 7: collisions, 8: velocities and friction
            const float4 pi = P[i];
            if (pi.w <= 0.0f) continue;
            const v3 pv = ld3(prv[i]);
            v3 p = ld3(pi);
            bool contacted = false;
            v3 cn = splat(0.0f);
            float cd = 0.0f;
            for (uint32_t c = 0; c < n_col; ++c) {
                const uint32_t slot = cols[c];
                const uint4 info = d.s_info[slot];
                const ShapeGeom sh = load_shape(d, slot, info);
                v3 nrm = splat(0.0f);
                float dep = 0.0f;
                if (project_particle(d, sh, ld3(d.b_pos[info.x]), ldq(d.b_rot[info.x]), r.thickness, p, pv, nrm, dep)) {
                    if (!contacted || dep > cd) {
                        cn = nrm;
                        cd = dep;
                    }
                    contacted = true;
                }
            }
            v3 v = (p - pv) * inv_hs;
            if (contacted && r.friction > 0.0f) {
                const v3 motion = p - pv;
                const v3 tang = motion - cn * dot(motion, cn);
                const float tl = len(tang);
                if (tl > 1e-6f) {
                    const float limit = r.friction * cd;
                    const float scale = gmax(0.0f, 1.0f - gmin(limit, tl) / tl);
                    p = p - tang * (1.0f - scale);
                    v = (p - pv) * inv_hs;
                }
            }
            vel[i] = st3(v, 0.0f);
            P[i] = st3(p, pi.w);
        }
        time += hs;
        __syncthreads();
    }
    for (uint32_t i = tid; i < r.np; i += kSoftBlock) pos[i] = P[i];
    v3 lo, hi;
    fold_bounds(P, r.np, part_lo, part_hi, lo, hi);
    float s2 = 0.0f;
    for (uint32_t i = tid; i < r.np; i += kSoftBlock) {
        const v3 v = ld3(vel[i]);
        s2 = fmaxf(s2, dot(v, v));
    }
    const float speed = sqrtf(block_maxf(s2, red));
    bool sleep_now = false;
    if (speed < r.sleep_threshold) {
        stt.sleep_timer += dt;
        if (stt.sleep_timer >= r.sleep_time) {
            stt.awake = 0u;
            sleep_now = true;
        }
    } else {
        stt.sleep_timer = 0.0f;
    }
    if (sleep_now)
        for (uint32_t i = tid; i < r.np; i += kSoftBlock) vel[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (tid == 0) {
        stt.time = time;
        stt.bmin[0] = lo.x; stt.bmin[1] = lo.y; stt.bmin[2] = lo.z;
        stt.bmax[0] = hi.x; stt.bmax[1] = hi.y; stt.bmax[2] = hi.z;
        stt.error = overflow ? 1u : 0u;
        s.st_state[si] = stt;
    }
}

// sync_broadphase_ (world.h:1511-1518) for a frame with no rigid substep: the dynamic proxies'
// fat boxes refit around their tight bounds swept by the velocity over h, as k_bounds does
__global__ void k_refit(DevState d, float h, float margin) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.ns) return;
    const uint4 info = d.s_info[i];
    if (!shape_live(d, info) || (info.y & SF_STATIC)) return;
    const ShapeGeom s = load_shape(d, i, info);
    v3 mn, mx;
    const uint32_t ob = info.x;
    shape_bounds(d, s, ld3(d.b_pos[ob]), ldq(d.b_rot[ob]), mn, mx);
    const v3 fmn = ld3(d.s_fmin[i]), fmx = ld3(d.s_fmax[i]);
    const v3 disp = ld3(d.b_vel[ob]) * h;
    const v3 smn = vmin(mn, mn + disp), smx = vmax(mx, mx + disp);
    const bool contains = smn.x >= fmn.x && smn.x <= fmx.x && smn.y >= fmn.y && smn.y <= fmx.y && smn.z >= fmn.z &&
                          smn.z <= fmx.z && smx.x >= fmn.x && smx.x <= fmx.x && smx.y >= fmn.y && smx.y <= fmx.y &&
                          smx.z >= fmn.z && smx.z <= fmx.z;
    if (!contains) {
        d.s_fmin[i] = st3(smn - splat(margin), 0.0f);
        d.s_fmax[i] = st3(smx + splat(margin), 0.0f);
    }
}

}  // namespace

void launch_refit(const DevState& d, float h, float margin, hipStream_t st) {
    if (d.ns) hipLaunchKernelGGL(k_refit, dim3((d.ns + 255) / 256), dim3(256), 0, st, d, h, margin);
}
void launch_cloths(const DevState& d, const SoftDev& s, uint32_t max_np, float dt, uint32_t frame_substeps,
                   uint32_t world_iters, v3 gravity, float margin, hipStream_t st) {
    if (!s.n_cloths) return;
    hipLaunchKernelGGL(k_cloth, dim3(s.n_cloths), dim3(kSoftBlock), size_t(max_np) * sizeof(float4), st, d, s, dt,
                       frame_substeps, world_iters, gravity, margin);
}
void launch_strands(const DevState& d, const SoftDev& s, uint32_t max_np, float dt, uint32_t frame_substeps,
                    uint32_t world_iters, v3 gravity, float margin, hipStream_t st) {
    if (!s.n_sets) return;
    hipLaunchKernelGGL(k_strands, dim3(s.n_sets), dim3(kSoftBlock), size_t(max_np) * sizeof(float4), st, d, s, dt,
                       frame_substeps, world_iters, gravity, margin);
}
hipError_t soft_lds_limit(size_t bytes) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_cloth), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(bytes));
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_strands), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(bytes));
}

}  // namespace pxg
