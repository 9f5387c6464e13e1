// Batched stepping of thousands of independent scenes in one persistent launch.
//
// One workgroup owns one scene for all K steps of a launch: the scene's bodies, broadphase boxes,
// pair list, colours and solve order live in LDS, so every phase of the step is a workgroup-local
// loop separated by __syncthreads() instead of kernel launches or grid barriers. Scenes are taken
// from a work queue in decreasing size (longest first) so uneven scenes balance across CUs. Manifolds
// (pair-indexed) and the warm-start cache live in per-scene regions of global memory (L2-resident
// while the scene is being stepped). The cache is an open-addressing table whose keys carry the step
// generation, so it is never cleared: an entry from an older step reads as empty.
//
// Per step, the same algorithm as the single-world path (and the reference's step_fixed,
// physxcoopa/world.h:1278-1317): forces, fat-AABB update, pairs (all pairs of the scene in (a, b)
// order), narrowphase, colouring, warm start, velocity/relax/position passes colour by colour,
// cache rebuild, islands and sleep, integration.
#include "../../include/pxg.h"
#include "pxg_device.h"
#include "pxg_joints.h"
#include "pxg_mlp.h"
#include "pxg_state.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

using namespace pxd;

namespace pxb {

constexpr int kMaxBodies = 288;   // bodies (and shapes) per scene, statics included
constexpr int kMaxPairs = 3072;   // broadphase pairs per scene per step
constexpr uint32_t kUncol = 0xFFu, kSkip = 0xFEu, kTrig = 0xFDu, kInvalid = 0xFCu;
constexpr int kMaxColoursB = 62;   // colours per scene (uint8 colour tags above 0xFB are reserved)
constexpr uint32_t SERR_PAIRS = 1u, SERR_COLOURS = 2u, SERR_CACHE = 4u, SERR_SIZE = 8u;

constexpr uint32_t BF_TYPE = 3u, BF_ALIVE = 1u << 2, BF_AWAKE = 1u << 3, BF_GRAVITY = 1u << 4;
constexpr uint32_t SF_ENABLED = 1u << 4, SF_TRIGGER = 1u << 5, SF_STATIC = 1u << 6;

struct Scene {
    uint32_t body_off, n_bodies, shape_off, n_shapes;
    uint32_t man_off, man_cap;     // pair-indexed manifold region
    uint32_t cache_off, cache_mask;
    uint32_t j_off, n_j;           // joints [j_off, j_off + n_j) (scene-local body indices)
    uint32_t j_layers, jl_off;     // layers of one joint pass; offsets at j_loff[jl_off .. + 2 (L + 1))
    uint32_t nc_off, n_nc;         // sorted body pairs without contacts, at nc_keys[nc_off ..]
};

struct SceneStats {
    uint32_t pairs, manifolds, points, colours;
    uint32_t active, passes, err, ordered;
    float residual;
    uint32_t awake, reserved[2];
};

struct BatchDev {
    uint32_t n_scenes;
    const Scene* scenes;
    const uint32_t* queue_order;  // scenes by decreasing size
    uint32_t* queue;              // work-queue counter
    SceneStats* stats;
    // bodies (concatenated)
    float4 *b_pos, *b_rot, *b_vel, *b_angv, *b_force, *b_torque, *b_iil, *b_iiw, *b_prev_pos, *b_prev_rot;
    uint32_t* b_flags;
    // shapes (concatenated; owner is scene-local)
    const uint4* s_info;
    const float4 *s_lc, *s_lr, *s_dim, *s_mat;
    float4 *s_fmin, *s_fmax;
    // manifolds (pair-indexed per scene region)
    uint32_t* m_valid;            // 1 = valid manifold
    uint2* m_pair;                // shape pair (scene-local)
    uint2* m_ab;                  // body pair (scene-local)
    float4* m_nf;                 // normal, friction
    uint4* m_meta;                // count | trigger << 8, restitution bits, colour, solve position
    float4* m_p;                  // 4 per manifold: position, penetration
    float4* m_q;                  // 4 per manifold: separation, feature id bits, restitution bias, miss flag
    float4* m_imp;                // 4 per manifold: impulses
    uint16_t* m_order;            // per scene region: solve order (pair indices)
    uint32_t* m_colour_off;       // per scene: 66 entries
    // warm-start cache (per scene region): key | generation << 50, order (generation << 32 | pos)
    unsigned long long* c_key;
    unsigned long long* c_ord;
    float4* c_val;
    const uint32_t* gen0;         // generation of the first step of this launch (device)
    const uint32_t* layer_mask;   // 32 rows
    // joints (pxg_joints.h records; type, a, b, live in j_meta); each scene's solve order is its
    // joints in layers, forward (j_order[2 j_off ..]) and reverse (j_order[2 j_off + n_j ..])
    const uint4* j_meta;
    const float4* j_def;
    float4* j_acc;
    const uint16_t* j_order;      // scene-local joint indices
    const uint32_t* j_loff;
    const uint32_t* nc_keys;      // lo << 16 | hi
    // learned warm start: network (nullptr = off) and per-scene staging of the warm-start misses of
    // the last step (kMissRows rows per scene: features, point index, clamped prediction)
    const pxg::MlpDev* mlp;
    float* feat;
    uint32_t* miss_pt;
    float* pred;
};
constexpr uint32_t kMissRows = 256;  // misses predicted per scene and step; the rest start at zero

struct Cfg {
    float h, gx, gy, gz, margin;
    float pos_corr, slop, max_corr, rest_thresh, sleep_lin, sleep_ang, sleep_time, max_lin_vel;
    uint32_t exact_quat;
    uint32_t vel_iters, relax_iters, pos_iters, residual_mode, residual_max;
    float residual_tol;
    uint32_t steps;
    uint32_t check_every;  // residual mode: the residual after every k-th pass and the last
};

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t hash64(unsigned long long k) {
    return hash32(uint32_t(k) ^ hash32(uint32_t(k >> 32) + 0x9e3779b9u));
}

// Shared-memory image of one scene.
struct Lds {
    float4 pos[kMaxBodies];    // position, sleep timer
    float4 rot[kMaxBodies];
    float4 vel[kMaxBodies];    // linear velocity, inverse mass
    float4 angv[kMaxBodies];   // angular velocity, mass
    float4 iiw[3 * kMaxBodies];
    uint32_t flags[kMaxBodies];
    float4 fmin[kMaxBodies], fmax[kMaxBodies], tmin[kMaxBodies], tmax[kMaxBodies];
    uint32_t pairs[kMaxPairs];          // a << 16 | b (shape indices)
    uint8_t colour[kMaxPairs];
    uint16_t order[kMaxPairs];
    unsigned long long best[kMaxBodies];
    unsigned long long mask[kMaxBodies];
    uint16_t parent[kMaxBodies];
    uint32_t isl_min[kMaxBodies];
    uint32_t scan[16];
    uint32_t colour_off[kMaxColoursB + 2];
    uint32_t n_pairs, remaining, colours, err, active, ordered, points, passes, awake, n_valid;
    float residual;
    uint32_t scene;
    uint32_t n_miss;                    // learned warm start: misses queued this step
    uint32_t passes_total;              // velocity passes summed over the steps of this launch
};
// The network's per-wave hidden-activation buffers (16 rows x 64 BF16 each) reuse the pair list,
// which is dead between the narrowphase and the next step's broadphase; at most kMlpWaves waves
// of a workgroup run the network.
constexpr uint32_t kMlpWaves = 4;
static_assert(kMlpWaves * 16 * pxg::kMlpH * sizeof(uint16_t) <= kMaxPairs * sizeof(uint32_t), "MLP buffers fit the pair list");
__device__ __forceinline__ uint16_t (*mlp_buffer(Lds& L, uint32_t wave))[pxg::kMlpH] {
    return reinterpret_cast<uint16_t (*)[pxg::kMlpH]>(reinterpret_cast<uint16_t*>(L.pairs) + wave * 16u * pxg::kMlpH);
}

}  // namespace pxb

namespace pxb {

__device__ __forceinline__ bool active_flag(uint32_t f) { return (f & BF_TYPE) == 2u && (f & BF_AWAKE) && (f & BF_ALIVE); }

__device__ __forceinline__ m3 iiw_of(const Lds& L, uint32_t i) {
    float4 a = L.iiw[3 * i], b = L.iiw[3 * i + 1], c = L.iiw[3 * i + 2];
    m3 m;
    m.c0 = mk(a.x, a.y, a.z);
    m.c1 = mk(a.w, b.x, b.y);
    m.c2 = mk(b.z, b.w, c.x);
    return m;
}
__device__ __forceinline__ void set_iiw(Lds& L, uint32_t i, const m3& m) {
    L.iiw[3 * i] = make_float4(m.c0.x, m.c0.y, m.c0.z, m.c1.x);
    L.iiw[3 * i + 1] = make_float4(m.c1.y, m.c1.z, m.c2.x, m.c2.y);
    L.iiw[3 * i + 2] = make_float4(m.c2.z, 0.0f, 0.0f, 0.0f);
}

struct LB {
    v3 v, w, x;
    float im;
    bool act;
    m3 I;
};
__device__ __forceinline__ LB lb_load(const Lds& L, uint32_t i) {
    LB b;
    uint32_t f = L.flags[i];
    b.act = active_flag(f);
    float4 v = L.vel[i], w = L.angv[i];
    b.v = ld3(v);
    b.w = ld3(w);
    b.im = b.act ? v.w : 0.0f;
    if (b.act) b.I = iiw_of(L, i);
    else { b.I.c0 = splat(0.0f); b.I.c1 = splat(0.0f); b.I.c2 = splat(0.0f); }
    b.x = ld3(L.pos[i]);
    return b;
}
__device__ __forceinline__ void lb_store(Lds& L, uint32_t i, const LB& b) {
    if (!b.act) return;
    L.vel[i] = st3(b.v, L.vel[i].w);
    L.angv[i] = st3(b.w, L.angv[i].w);
}
// apply_impulse_pair (dynamics/solver.h:86-95)
__device__ __forceinline__ void lapply(LB& a, LB& b, v3 p, v3 J) {
    if (a.act) { a.v = a.v - J * a.im; a.w = a.w - mul(a.I, cross(p - a.x, J)); }
    if (b.act) { b.v = b.v + J * b.im; b.w = b.w + mul(b.I, cross(p - b.x, J)); }
}

// joint rows on the scene's LDS bodies
__device__ __forceinline__ JBody jb_lds(const Lds& L, uint32_t i) {
    JBody b;
    b.act = active_flag(L.flags[i]);
    float4 v = L.vel[i], w = L.angv[i];
    b.v = ld3(v);
    b.w = ld3(w);
    b.im = b.act ? v.w : 0.0f;
    b.I = b.act ? iiw_of(L, i) : zero3();
    b.x = ld3(L.pos[i]);
    b.q = ldq(L.rot[i]);
    return b;
}
__device__ __forceinline__ void jb_store_vel(Lds& L, uint32_t i, const JBody& b) {
    if (!b.act) return;
    L.vel[i] = st3(b.v, L.vel[i].w);
    L.angv[i] = st3(b.w, L.angv[i].w);
}
__device__ __forceinline__ void jb_store_pose(Lds& L, uint32_t i, const JBody& b) {
    if (!b.act) return;
    L.pos[i] = st3(b.x, L.pos[i].w);
    L.rot[i] = stq(b.q);
}
// body pairs without contacts (world.h:1619-1620): binary search of the scene's sorted keys
__device__ __forceinline__ bool scene_contacts_off(const BatchDev& bd, const Scene& S, uint32_t a, uint32_t b) {
    if (!S.n_nc) return false;
    const uint32_t key = ((a < b ? a : b) << 16) | (a < b ? b : a);
    const uint32_t* k = bd.nc_keys + S.nc_off;
    uint32_t lo = 0, hi = S.n_nc;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (k[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo < S.n_nc && k[lo] == key;
}
// one joint pass over the scene's layers (mode 0 warm start, 1 velocity, 2 position), block
// barriers between layers
template <int TB>
__device__ __forceinline__ void scene_joint_pass(const BatchDev& bd, const Cfg& cfg, Lds& L, const Scene& S, int mode,
                                                 bool rev) {
    const uint32_t* off = bd.j_loff + S.jl_off + (rev ? S.j_layers + 1 : 0u);
    const uint16_t* ord = bd.j_order + 2u * S.j_off + (rev ? S.n_j : 0u);
    const float inv_h = cfg.h > 0.0f ? 1.0f / cfg.h : 0.0f;
    for (uint32_t l = 0; l < S.j_layers; ++l) {
        for (uint32_t t = off[l] + threadIdx.x; t < off[l + 1]; t += TB) {
            const uint32_t j = S.j_off + ord[t];
            const uint4 m = bd.j_meta[j];
            if (!m.w) continue;
            JDef def = load_jdef(bd.j_def, j);
            JBody a = jb_lds(L, m.y), b = jb_lds(L, m.z);
            if (mode == 0) {
                JAcc acc = load_jacc(bd.j_acc, j);
                acc.limit = 0.0f;
                acc.swing = 0.0f;
                joint_warm(def, acc, a, b);
                store_jacc(bd.j_acc, j, acc);
                jb_store_vel(L, m.y, a);
                jb_store_vel(L, m.z, b);
            } else if (mode == 1) {
                if (!a.act && !b.act) continue;
                JAcc acc = load_jacc(bd.j_acc, j);
                joint_velocity(def, acc, a, b, inv_h);
                store_jacc(bd.j_acc, j, acc);
                jb_store_vel(L, m.y, a);
                jb_store_vel(L, m.z, b);
            } else {
                if (!a.act && !b.act) continue;
                joint_position(def, a, b, cfg.slop, cfg.pos_corr, cfg.max_corr);
                jb_store_pose(L, m.y, a);
                jb_store_pose(L, m.z, b);
            }
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint32_t wsum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// exclusive scan across the workgroup (TB threads); all threads must call
template <int TB>
__device__ __forceinline__ uint32_t bscan(Lds& L, uint32_t v, uint32_t& total) {
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    __syncthreads();
    if (lane == 63) L.scan[wid] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < TB / 64; ++w) {
        uint32_t s = L.scan[w];
        if (w < wid) off += s;
        tot += s;
    }
    total = tot;
    return off + x - v;
}

__device__ __forceinline__ ShapeGeom shape_geom(const BatchDev& bd, uint32_t gi) {
    uint4 info = bd.s_info[gi];
    ShapeGeom s;
    s.type = int(info.y & 15u);
    s.caxis = int((info.y >> 8) & 3u);
    float4 lc = bd.s_lc[gi];
    s.lc = mk(lc.x, lc.y, lc.z);
    s.radius = lc.w;
    s.lr = ldq(bd.s_lr[gi]);
    float4 dim = bd.s_dim[gi];
    s.he = mk(dim.x, dim.y, dim.z);
    s.cr = dim.w;
    s.chh = bd.s_mat[gi].x;
    return s;
}
// Contact generation is called out of line: the narrowphase's working set (clipping polygons,
// the manifold under construction) then lives only in this call's frame instead of adding to the
// register and scratch budget of the whole scene kernel.
__device__ __noinline__ bool scene_contacts(const ShapeGeom& si, v3 pi, quat qi, const ShapeGeom& sj, v3 pj, quat qj,
                                            bool spec, Man& m) {
    return generate_contacts(si, pi, qi, sj, pj, qj, spec, m);
}
__device__ __forceinline__ bool ovl(float4 amn, float4 amx, float4 bmn, float4 bmx) {
    return amn.x <= bmx.x && amx.x >= bmn.x && amn.y <= bmx.y && amx.y >= bmn.y && amn.z <= bmx.z && amx.z >= bmn.z;
}

// ---- one step of one scene; every thread of the block calls it ----
template <int TB>
__device__ void scene_step(const BatchDev& bd, const Cfg& cfg, Lds& L, const Scene& S, uint32_t gen, bool last,
                           const uint32_t* layer_mask) {
    const uint32_t t = threadIdx.x;
    const uint32_t nb = S.n_bodies, ns = S.n_shapes;
    // 1. gravity + force integration (integrator.h:31-35, 78-94)
    for (uint32_t i = t; i < nb; i += TB) {
        uint32_t f = L.flags[i];
        if (!(f & BF_ALIVE)) continue;
        float4 fo = bd.b_force[S.body_off + i], to = bd.b_torque[S.body_off + i];
        bool dyn = (f & BF_TYPE) == 2u, awake = (f & BF_AWAKE) != 0;
        if (dyn && awake) {
            float4 vm = L.vel[i], wm = L.angv[i];
            v3 force = ld3(fo), torque = ld3(to);
            if (f & BF_GRAVITY) force = force + mk(cfg.gx, cfg.gy, cfg.gz) * wm.w;
            v3 v = ld3(vm) + (force * vm.w) * cfg.h;
            v3 w = ld3(wm) + mul(iiw_of(L, i), torque) * cfg.h;
            v = v * (1.0f / (1.0f + cfg.h * fo.w));
            w = w * (1.0f / (1.0f + cfg.h * to.w));
            uint32_t cons = f >> 8;
            if (cons & 1u) v.x = 0.0f;
            if (cons & 2u) v.y = 0.0f;
            if (cons & 4u) v.z = 0.0f;
            if (cons & 8u) w.x = 0.0f;
            if (cons & 16u) w.y = 0.0f;
            if (cons & 32u) w.z = 0.0f;
            L.vel[i] = st3(v, vm.w);
            L.angv[i] = st3(w, wm.w);
        }
        if (fo.x != 0.0f || fo.y != 0.0f || fo.z != 0.0f) bd.b_force[S.body_off + i] = make_float4(0.0f, 0.0f, 0.0f, fo.w);
        if (to.x != 0.0f || to.y != 0.0f || to.z != 0.0f) bd.b_torque[S.body_off + i] = make_float4(0.0f, 0.0f, 0.0f, to.w);
    }
    __syncthreads();
    // 2. tight bounds and fat-AABB update (world.h:1489-1518, aabb_tree.h:76-85)
    for (uint32_t j = t; j < ns; j += TB) {
        uint4 info = bd.s_info[S.shape_off + j];
        uint32_t ob = info.x;
        if (!(info.y & SF_ENABLED) || ob >= nb || !(L.flags[ob] & BF_ALIVE)) {
            L.tmin[j] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
            continue;
        }
        ShapeGeom s = shape_geom(bd, S.shape_off + j);
        v3 mn, mx;
        world_bounds(s, ld3(L.pos[ob]), ldq(L.rot[ob]), mn, mx);
        float4 fmn = L.fmin[j], fmx = L.fmax[j];
        if (!(info.y & SF_STATIC)) {
            v3 disp = ld3(L.vel[ob]) * cfg.h;
            v3 smn = vmin(mn, mn + disp), smx = vmax(mx, mx + disp);
            bool contains = smn.x >= fmn.x && smn.x <= fmx.x && smn.y >= fmn.y && smn.y <= fmx.y && smn.z >= fmn.z &&
                            smn.z <= fmx.z && smx.x >= fmn.x && smx.x <= fmx.x && smx.y >= fmn.y && smx.y <= fmx.y &&
                            smx.z >= fmn.z && smx.z <= fmx.z;
            if (!contains) {
                L.fmin[j] = st3(smn - splat(cfg.margin), 0.0f);
                L.fmax[j] = st3(smx + splat(cfg.margin), 0.0f);
            }
        }
        L.tmin[j] = st3(mn, (info.y & SF_STATIC) ? 1.0f : 0.0f);
        L.tmax[j] = st3(mx, __uint_as_float(ob));
    }
    __syncthreads();
    // 3. pairs in (a, b) order: thread t owns rows [t*R, t*R+R) (world.h:1545-1577 pair rule)
    const uint32_t R = (ns + TB - 1) / TB;
    auto row_pairs = [&](uint32_t a, bool write, uint32_t& n) {
        float4 ta = L.tmin[a];
        if (ta.w < 0.0f) return;
        float4 tA = L.tmax[a], fa = L.fmin[a], fA = L.fmax[a];
        uint32_t oa = __float_as_uint(tA.w);
        for (uint32_t b = a + 1; b < ns; ++b) {
            float4 tb = L.tmin[b];
            if (tb.w < 0.0f) continue;
            float4 tB = L.tmax[b];
            if (__float_as_uint(tB.w) == oa) continue;
            if (ta.w > 0.0f && tb.w > 0.0f) continue;  // both static
            if (ovl(ta, tA, L.fmin[b], L.fmax[b]) || ovl(tb, tB, fa, fA)) {
                if (write && n < uint32_t(kMaxPairs)) L.pairs[n] = (a << 16) | b;
                ++n;
            }
        }
    };
    uint32_t cnt = 0;
    for (uint32_t a = t * R; a < t * R + R && a < ns; ++a) row_pairs(a, false, cnt);
    uint32_t np = 0;
    uint32_t off = bscan<TB>(L, cnt, np);
    for (uint32_t a = t * R; a < t * R + R && a < ns; ++a) row_pairs(a, true, off);
    if (np > uint32_t(kMaxPairs) || np > S.man_cap) {
        if (t == 0) L.err |= SERR_PAIRS;
        np = min(uint32_t(kMaxPairs), S.man_cap);
    }
    if (t == 0) L.n_pairs = np;
    __syncthreads();
    // wake_jointed_islands (solver.h:887-901), the first act of solve(): before the manifolds'
    // activity is classified below. Sweeps until nothing changes.
    for (uint32_t pass = 0; S.n_j && pass <= S.n_j; ++pass) {
        int changed = 0;
        for (uint32_t j = t; j < S.n_j; j += TB) {
            const uint4 m = bd.j_meta[S.j_off + j];
            if (!m.w) continue;
            uint32_t fa = L.flags[m.y], fb = L.flags[m.z];
            if ((fa & BF_TYPE) != 2u || (fb & BF_TYPE) != 2u) continue;
            bool aw = (fa & BF_AWAKE) != 0u, bw = (fb & BF_AWAKE) != 0u;
            if (aw == bw) continue;
            uint32_t sl = aw ? m.z : m.y;
            atomicOr(&L.flags[sl], BF_AWAKE);
            L.pos[sl].w = 0.0f;
            changed = 1;
        }
        if (!__syncthreads_or(changed)) break;
    }
    // 4. narrowphase (world.h:1598-1654), pair-indexed manifolds
    uint32_t pts = 0;
    for (uint32_t k = t; k < np; k += TB) {
        uint32_t pr = L.pairs[k], i = pr >> 16, j = pr & 0xFFFFu;
        uint4 ii = bd.s_info[S.shape_off + i], ij = bd.s_info[S.shape_off + j];
        uint32_t bi = ii.x, bj = ij.x;
        uint32_t fi = L.flags[bi], fj = L.flags[bj];
        uint32_t gm = S.man_off + k;
        Man m;
        m.count = 0;
        bool valid = false;
        if ((fi & BF_ALIVE) && (fj & BF_ALIVE) && ((fi & BF_TYPE) == 2u || (fj & BF_TYPE) == 2u) &&
            (layer_mask[ii.z & 31u] & (1u << (ij.z & 31u))) && !scene_contacts_off(bd, S, bi, bj)) {
            ShapeGeom si = shape_geom(bd, S.shape_off + i), sj = shape_geom(bd, S.shape_off + j);
            bool spec = (len(ld3(L.vel[bi])) * cfg.h > min_extent(si)) || (len(ld3(L.vel[bj])) * cfg.h > min_extent(sj));
            valid = scene_contacts(si, ld3(L.pos[bi]), ldq(L.rot[bi]), sj, ld3(L.pos[bj]), ldq(L.rot[bj]), spec, m);
        }
        bd.m_valid[gm] = valid ? 1u : 0u;
        L.colour[k] = kInvalid;
        if (!valid) continue;
        float4 mi = bd.s_mat[S.shape_off + i], mj = bd.s_mat[S.shape_off + j];
        float fr = combine(mi.y, mj.y, int(ii.w & 255u), int(ij.w & 255u));
        float re = combine(mi.z, mj.z, int(ii.w >> 8), int(ij.w >> 8));
        uint32_t trig = ((ii.y | ij.y) & SF_TRIGGER) ? 1u : 0u;
        bd.m_pair[gm] = make_uint2(i, j);
        bd.m_ab[gm] = make_uint2(bi, bj);
        bd.m_nf[gm] = st3(m.n, fr);
        bd.m_meta[gm] = make_uint4(uint32_t(m.count) | (trig << 8), __float_as_uint(re), 0u, 0u);
        for (int q = 0; q < 4; ++q) {
            if (q < m.count) {
                bd.m_p[4 * gm + q] = st3(m.p[q].pos, m.p[q].pen);
                bd.m_q[4 * gm + q] = make_float4(m.p[q].sep, __uint_as_float(m.p[q].fid), 0.0f, 0.0f);
            } else {
                bd.m_p[4 * gm + q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                bd.m_q[4 * gm + q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
        pts += uint32_t(m.count);
        bool act = active_flag(fi) || active_flag(fj);
        L.colour[k] = trig ? kTrig : (act ? kUncol : kSkip);
    }
    // 5. colouring (Jones-Plassmann rounds, first-fit colour; only awake dynamic bodies constrain)
    for (uint32_t i = t; i < nb; i += TB) { L.best[i] = 0ull; L.mask[i] = 0ull; }
    if (t == 0) { L.remaining = 0; L.colours = 0; L.points = 0; L.n_valid = 0; L.n_miss = 0; }
    __syncthreads();
    uint32_t rem = 0, nval = 0;
    for (uint32_t k = t; k < np; k += TB) {
        rem += L.colour[k] == kUncol ? 1u : 0u;
        nval += L.colour[k] != kInvalid ? 1u : 0u;
    }
    rem = wsum(rem);
    pts = wsum(pts);
    nval = wsum(nval);
    if ((t & 63) == 0) { atomicAdd(&L.remaining, rem); atomicAdd(&L.points, pts); atomicAdd(&L.n_valid, nval); }
    __syncthreads();
    for (uint32_t round = 1; round < 1024u; ++round) {
        if (L.remaining == 0u) break;
        for (uint32_t k = t; k < np; k += TB) {
            if (L.colour[k] != kUncol) continue;
            uint2 ab = bd.m_ab[S.man_off + k];
            uint32_t pr = hash32(ab.x * 0x9E3779B1u ^ hash32(ab.y + 0x85EBCA77u)) & 0xFFFu;
            unsigned long long key = (static_cast<unsigned long long>(round) << 44) | (static_cast<unsigned long long>(pr) << 32) | k;
            if (active_flag(L.flags[ab.x])) atomicMax(&L.best[ab.x], key);
            if (active_flag(L.flags[ab.y])) atomicMax(&L.best[ab.y], key);
        }
        __syncthreads();
        for (uint32_t k = t; k < np; k += TB) {
            if (L.colour[k] != kUncol) continue;
            uint2 ab = bd.m_ab[S.man_off + k];
            uint32_t pr = hash32(ab.x * 0x9E3779B1u ^ hash32(ab.y + 0x85EBCA77u)) & 0xFFFu;
            unsigned long long key = (static_cast<unsigned long long>(round) << 44) | (static_cast<unsigned long long>(pr) << 32) | k;
            bool aa = active_flag(L.flags[ab.x]), ba = active_flag(L.flags[ab.y]);
            if ((aa && L.best[ab.x] != key) || (ba && L.best[ab.y] != key)) continue;
            unsigned long long freeb = ~((aa ? L.mask[ab.x] : 0ull) | (ba ? L.mask[ab.y] : 0ull));
            if (freeb == 0ull || (freeb & ((1ull << kMaxColoursB) - 1ull)) == 0ull) { atomicOr(&L.err, SERR_COLOURS); continue; }
            uint32_t c = uint32_t(__ffsll(static_cast<long long>(freeb)) - 1);
            L.colour[k] = uint8_t(c);
            if (aa) atomicOr(&L.mask[ab.x], 1ull << c);
            if (ba) atomicOr(&L.mask[ab.y], 1ull << c);
            atomicSub(&L.remaining, 1u);
            atomicMax(&L.colours, c + 1u);
        }
        __syncthreads();
        if (L.err & SERR_COLOURS) break;
    }
    __syncthreads();
    // 6. solve order: colours in order, then the skipped non-trigger manifolds; stable in pair order
    const uint32_t C = L.colours;
    const uint32_t Rp = (np + TB - 1) / TB;
    uint32_t base = 0;
    for (uint32_t b = 0; b <= C; ++b) {
        uint32_t want = b < C ? b : kSkip;
        uint32_t c0 = 0;
        for (uint32_t k = t * Rp; k < t * Rp + Rp && k < np; ++k) c0 += L.colour[k] == want ? 1u : 0u;
        uint32_t tot = 0;
        uint32_t o = bscan<TB>(L, c0, tot) + base;
        for (uint32_t k = t * Rp; k < t * Rp + Rp && k < np; ++k)
            if (L.colour[k] == want) L.order[o++] = uint16_t(k);
        if (t == 0) L.colour_off[b] = base;
        base += tot;
    }
    if (t == 0) { L.colour_off[C + 1] = base; L.active = L.colour_off[C]; L.ordered = base; }
    __syncthreads();
    const uint32_t active = L.active, ordered = L.ordered;
    // 7. restitution targets (solver.h:121-137) and warm-start lookup (solver.h:144-175)
    const uint32_t cmask = S.cache_mask;
    for (uint32_t s = t; s < ordered; s += TB) {
        uint32_t k = L.order[s], gm = S.man_off + k;
        uint2 ab = bd.m_ab[gm];
        float4 nf = bd.m_nf[gm];
        uint4 meta = bd.m_meta[gm];
        uint32_t cnt = meta.x & 255u;
        float rest = __uint_as_float(meta.y);
        v3 n = ld3(nf);
        v3 pa = ld3(L.pos[ab.x]), pb = ld3(L.pos[ab.y]);
        v3 va = ld3(L.vel[ab.x]), vb = ld3(L.vel[ab.y]);
        v3 wa = ld3(L.angv[ab.x]), wb = ld3(L.angv[ab.y]);
        for (uint32_t q = 0; q < cnt; ++q) {
            float4 pp = bd.m_p[4 * gm + q], qq = bd.m_q[4 * gm + q];
            v3 pos = ld3(pp);
            v3 ra = pos - pa, rb = pos - pb;
            float vap = dot((vb + cross(wb, rb)) - (va + cross(wa, ra)), n);
            qq.z = (-vap > cfg.rest_thresh) ? -rest * vap : 0.0f;
            float4 imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            qq.w = 0.0f;
            if (s < active) {
                unsigned long long key = pack_key(ab.x, ab.y, __float_as_uint(qq.y)) | (static_cast<unsigned long long>(gen - 1u) << 50);
                uint32_t idx = hash64(key) & cmask;
                bool hit = false;
                for (uint32_t probe = 0; probe <= cmask; ++probe) {
                    unsigned long long v = bd.c_key[S.cache_off + idx];
                    if (v == key) { imp = bd.c_val[S.cache_off + idx]; hit = true; break; }
                    if ((v >> 50) != ((gen - 1u) & 0x3FFFu)) break;
                    idx = (idx + 1u) & cmask;
                }
                if (!hit) {
                    qq.w = 1.0f;
                    if (bd.mlp) {  // queue the miss for the network (features of the pre-solve state)
                        uint32_t row = atomicAdd(&L.n_miss, 1u);
                        if (row < kMissRows) {
                            pxg::FeatBody FA, FB;
                            const bool aa = active_flag(L.flags[ab.x]), ba = active_flag(L.flags[ab.y]);
                            FA.x = pa; FA.v = va; FA.w = wa; FA.im = aa ? L.vel[ab.x].w : 0.0f;
                            FA.I = aa ? iiw_of(L, ab.x) : zero3();
                            FB.x = pb; FB.v = vb; FB.w = wb; FB.im = ba ? L.vel[ab.y].w : 0.0f;
                            FB.I = ba ? iiw_of(L, ab.y) : zero3();
                            float f[pxg::kMlpIn];
                            pxg::contact_features(FA, FB, n, pos, pp.w, qq.x, nf.w, rest, f);
                            const size_t r0 = size_t(L.scene) * kMissRows + row;
                            for (int c = 0; c < pxg::kMlpIn; ++c) bd.feat[r0 * pxg::kMlpIn + c] = f[c];
                            bd.miss_pt[r0] = 4u * gm + q;
                        }
                    }
                }
            }
            bd.m_q[4 * gm + q] = qq;
            bd.m_imp[4 * gm + q] = imp;
        }
    }
    __syncthreads();
    if (bd.mlp) {
        // the learned warm start: the network on the matrix cores, 16 misses per wave at a time,
        // then each prediction clamped to the friction cone becomes the point's initial impulse
        const uint32_t nm = L.n_miss < kMissRows ? L.n_miss : kMissRows;
        const size_t base = size_t(L.scene) * kMissRows;
        const uint32_t wave = t >> 6;
        const uint32_t nwaves = (TB / 64) < kMlpWaves ? (TB / 64) : kMlpWaves;
        const pxg::MlpView W = pxg::mlp_view(bd.mlp);
        if (wave < nwaves)
            for (uint32_t row0 = wave * 16u; row0 < nm; row0 += nwaves * 16u)
                pxg::mlp_rows16(W, bd.feat + base * pxg::kMlpIn, row0, nm, mlp_buffer(L, wave), bd.pred + base * pxg::kMlpOut);
        __syncthreads();
        for (uint32_t r = t; r < nm; r += TB) {
            const uint32_t pt = bd.miss_pt[base + r];
            const float mu = bd.m_nf[pt / 4u].w;
            float* pr = bd.pred + (base + r) * pxg::kMlpOut;
            const float ln = fmaxf(pr[0], 0.0f), lim = mu * ln;
            const float t1 = fminf(fmaxf(pr[1], -lim), lim), t2 = fminf(fmaxf(pr[2], -lim), lim);
            pr[0] = ln; pr[1] = t1; pr[2] = t2;
            bd.m_imp[pt] = make_float4(ln, t1, t2, 0.0f);
        }
        __syncthreads();
    }
    // 8. warm start, colour by colour
    for (uint32_t c = 0; c < C; ++c) {
        for (uint32_t s = L.colour_off[c] + t; s < L.colour_off[c + 1]; s += TB) {
            uint32_t gm = S.man_off + L.order[s];
            uint2 ab = bd.m_ab[gm];
            LB A = lb_load(L, ab.x), B = lb_load(L, ab.y);
            v3 n = ld3(bd.m_nf[gm]);
            uint32_t cnt = bd.m_meta[gm].x & 255u;
            v3 t1, t2;
            orthonormal_basis(n, t1, t2);
            for (uint32_t q = 0; q < cnt; ++q) {
                float4 imp = bd.m_imp[4 * gm + q];
                v3 J = n * imp.x + t1 * imp.y + t2 * imp.z;
                if (dot(J, J) > 0.0f) lapply(A, B, ld3(bd.m_p[4 * gm + q]), J);
            }
            lb_store(L, ab.x, A);
            lb_store(L, ab.y, B);
        }
        __syncthreads();
    }
    if (S.n_j) scene_joint_pass<TB>(bd, cfg, L, S, 0, false);  // warm_start_joints (solver.h:917)
    // 9. velocity and relax passes (solver.h:212-281), each followed by a joint pass alternating
    // direction (solver.h:939-946)
    auto vpass = [&](int use_rest, uint32_t offset) {
        for (uint32_t c = 0; c < C; ++c) {
            for (uint32_t s = L.colour_off[c] + t; s < L.colour_off[c + 1]; s += TB) {
                uint32_t gm = S.man_off + L.order[s];
                uint2 ab = bd.m_ab[gm];
                LB A = lb_load(L, ab.x), B = lb_load(L, ab.y);
                float4 nf = bd.m_nf[gm];
                v3 n = ld3(nf);
                uint32_t cnt = bd.m_meta[gm].x & 255u;
                v3 t1, t2;
                orthonormal_basis(n, t1, t2);
                for (uint32_t kk = 0; kk < cnt; ++kk) {
                    uint32_t q = (kk + offset) % cnt;
                    float4 pp = bd.m_p[4 * gm + q], qq = bd.m_q[4 * gm + q], imp = bd.m_imp[4 * gm + q];
                    v3 pos = ld3(pp);
                    v3 ra = pos - A.x, rb = pos - B.x;
                    {
                        v3 ran = cross(ra, n), rbn = cross(rb, n);
                        float kn = A.im + B.im + dot(ran, mul(A.I, ran)) + dot(rbn, mul(B.I, rbn));
                        if (kn > kEps) {
                            v3 va = A.v + cross(A.w, ra), vb = B.v + cross(B.w, rb);
                            float vn = dot(vb - va, n);
                            float target = qq.x > 0.0f ? -qq.x / cfg.h : (use_rest ? qq.z : 0.0f);
                            float lambda = -(vn - target) / kn;
                            float ni = gmax(imp.x + lambda, 0.0f);
                            float delta = ni - imp.x;
                            imp.x = ni;
                            lapply(A, B, pos, n * delta);
                        }
                    }
                    for (int tt = 0; tt < 2; ++tt) {
                        v3 dir = tt == 0 ? t1 : t2;
                        v3 rat = cross(ra, dir), rbt = cross(rb, dir);
                        float kt = A.im + B.im + dot(rat, mul(A.I, rat)) + dot(rbt, mul(B.I, rbt));
                        if (kt <= kEps) continue;
                        v3 va = A.v + cross(A.w, ra), vb = B.v + cross(B.w, rb);
                        float vt = dot(vb - va, dir);
                        float lt = -vt / kt;
                        float maxf = nf.w * imp.x;
                        float old = tt == 0 ? imp.y : imp.z;
                        float nt = sclamp(old + lt, -maxf, maxf);
                        float delta = nt - old;
                        if (tt == 0) imp.y = nt; else imp.z = nt;
                        lapply(A, B, pos, dir * delta);
                    }
                    bd.m_imp[4 * gm + q] = imp;
                }
                lb_store(L, ab.x, A);
                lb_store(L, ab.y, B);
            }
            __syncthreads();
        }
        if (S.n_j) scene_joint_pass<TB>(bd, cfg, L, S, 1, (offset & 1u) != 0u);
    };
    uint32_t passes = 0;
    float resid = 0.0f;
    uint32_t max_pass = cfg.residual_mode ? cfg.residual_max : cfg.vel_iters;
    for (uint32_t it = 0; it < max_pass; ++it) {
        vpass(0, it);
        ++passes;
        if (cfg.residual_mode && (passes % cfg.check_every == 0u || passes == max_pass)) {
            if (t == 0) L.residual = 0.0f;
            __syncthreads();
            float r = 0.0f;
            for (uint32_t s = t; s < active; s += TB) {
                uint32_t gm = S.man_off + L.order[s];
                uint2 ab = bd.m_ab[gm];
                LB A = lb_load(L, ab.x), B = lb_load(L, ab.y);
                float4 nf = bd.m_nf[gm];
                v3 n = ld3(nf);
                v3 t1, t2;
                orthonormal_basis(n, t1, t2);
                uint32_t cnt = bd.m_meta[gm].x & 255u;
                for (uint32_t q = 0; q < cnt; ++q) {
                    float4 pp = bd.m_p[4 * gm + q], qq = bd.m_q[4 * gm + q], imp = bd.m_imp[4 * gm + q];
                    v3 pos = ld3(pp);
                    v3 dv = (B.v + cross(B.w, pos - B.x)) - (A.v + cross(A.w, pos - A.x));
                    float vn = dot(dv, n);
                    float target = qq.x > 0.0f ? -qq.x / cfg.h : 0.0f;
                    r = gmax(r, imp.x > 0.0f ? fabsf(vn - target) : gmax(0.0f, target - vn));
                    float maxf = nf.w * imp.x;
                    for (int tt = 0; tt < 2; ++tt) {
                        float vt = dot(dv, tt == 0 ? t1 : t2);
                        float lt = tt == 0 ? imp.y : imp.z;
                        float rt = 0.0f;
                        if (maxf > 0.0f) rt = fabsf(lt) < maxf ? fabsf(vt) : gmax(0.0f, lt < 0.0f ? -vt : vt);
                        r = gmax(r, rt);
                    }
                }
            }
            for (int o = 32; o > 0; o >>= 1) r = fmaxf(r, __shfl_xor(r, o, 64));
            if ((t & 63) == 0 && r > 0.0f) atomicMax(reinterpret_cast<uint32_t*>(&L.residual), __float_as_uint(r));
            __syncthreads();
            resid = L.residual;
            __syncthreads();
            if (resid <= cfg.residual_tol) break;
        }
    }
    for (uint32_t it = 0; it < cfg.relax_iters; ++it) vpass(1, it);
    // 10. position passes (solver.h:293-328)
    for (uint32_t i = t; i < nb; i += TB) L.tmin[i] = L.pos[i];
    __syncthreads();
    for (uint32_t it = 0; it < cfg.pos_iters; ++it) {
        for (uint32_t c = 0; c < C; ++c) {
            for (uint32_t s = L.colour_off[c] + t; s < L.colour_off[c + 1]; s += TB) {
                uint32_t gm = S.man_off + L.order[s];
                uint2 ab = bd.m_ab[gm];
                bool aa = active_flag(L.flags[ab.x]), ba = active_flag(L.flags[ab.y]);
                float ima = aa ? L.vel[ab.x].w : 0.0f, imb = ba ? L.vel[ab.y].w : 0.0f;
                float kk = ima + imb;
                if (kk <= kEps) continue;
                float4 xa4 = L.pos[ab.x], xb4 = L.pos[ab.y];
                v3 xa = ld3(xa4), xb = ld3(xb4);
                v3 ra = ld3(L.tmin[ab.x]), rb = ld3(L.tmin[ab.y]);
                v3 n = ld3(bd.m_nf[gm]);
                v3 da = xa - ra, db = xb - rb;
                uint32_t cnt = bd.m_meta[gm].x & 255u;
                for (uint32_t q = 0; q < cnt; ++q) {
                    float pen = bd.m_p[4 * gm + q].w;
                    float cur = pen - dot(db - da, n);
                    float mag = gmax(0.0f, cur - cfg.slop) * cfg.pos_corr;
                    mag = gmin(mag, cfg.max_corr);
                    if (mag <= 0.0f) continue;
                    v3 corr = n * (mag / kk);
                    if (aa) { xa = xa - corr * ima; da = xa - ra; }
                    if (ba) { xb = xb + corr * imb; db = xb - rb; }
                }
                if (aa) L.pos[ab.x] = st3(xa, xa4.w);
                if (ba) L.pos[ab.y] = st3(xb, xb4.w);
            }
            __syncthreads();
        }
    }
    if (S.n_j) scene_joint_pass<TB>(bd, cfg, L, S, 2, false);  // solve_joint_position (solver.h:954)
    // 11. warm-start cache rebuild (solver.h:767-781): on a key collision the last writer in the
    // reference's manifold order (pair order, which is the pair index here) wins
    for (uint32_t s = t; s < ordered; s += TB) {
        uint32_t gm = S.man_off + L.order[s];
        uint2 ab = bd.m_ab[gm];
        uint32_t cnt = bd.m_meta[gm].x & 255u;
        for (uint32_t q = 0; q < cnt; ++q) {
            unsigned long long key = pack_key(ab.x, ab.y, __float_as_uint(bd.m_q[4 * gm + q].y)) |
                                     (static_cast<unsigned long long>(gen & 0x3FFFu) << 50);
            unsigned long long ord = (static_cast<unsigned long long>(gen) << 32) | (4u * L.order[s] + q + 1u);
            uint32_t idx = hash64(key) & cmask;
            bool done = false;
            for (uint32_t probe = 0; probe <= cmask && !done; ++probe) {
                unsigned long long* slot = &bd.c_key[S.cache_off + idx];
                unsigned long long v = *slot;
                while (true) {
                    if (v == key) { atomicMax(&bd.c_ord[S.cache_off + idx], ord); done = true; break; }
                    if ((v >> 50) == (gen & 0x3FFFu)) break;  // a current entry of another key
                    unsigned long long old = atomicCAS(slot, v, key);
                    if (old == v) { atomicMax(&bd.c_ord[S.cache_off + idx], ord); done = true; break; }
                    v = old;
                }
                idx = (idx + 1u) & cmask;
            }
            if (!done) atomicOr(&L.err, SERR_CACHE);
        }
    }
    __syncthreads();
    for (uint32_t s = t; s < ordered; s += TB) {
        uint32_t gm = S.man_off + L.order[s];
        uint2 ab = bd.m_ab[gm];
        uint32_t cnt = bd.m_meta[gm].x & 255u;
        for (uint32_t q = 0; q < cnt; ++q) {
            unsigned long long key = pack_key(ab.x, ab.y, __float_as_uint(bd.m_q[4 * gm + q].y)) |
                                     (static_cast<unsigned long long>(gen & 0x3FFFu) << 50);
            unsigned long long ord = (static_cast<unsigned long long>(gen) << 32) | (4u * L.order[s] + q + 1u);
            uint32_t idx = hash64(key) & cmask;
            for (uint32_t probe = 0; probe <= cmask; ++probe) {
                unsigned long long v = bd.c_key[S.cache_off + idx];
                if (v == key) {
                    if (bd.c_ord[S.cache_off + idx] == ord) bd.c_val[S.cache_off + idx] = bd.m_imp[4 * gm + q];
                    break;
                }
                if ((v >> 50) != (gen & 0x3FFFu)) break;
                idx = (idx + 1u) & cmask;
            }
        }
    }
    // 12. islands and sleep (solver.h:808-876, island.h:25-49)
    for (uint32_t i = t; i < nb; i += TB) { L.parent[i] = uint16_t(i); L.isl_min[i] = 0x7F800000u; }
    __syncthreads();
    for (uint32_t k = t; k < np; k += TB) {
        if (L.colour[k] == kInvalid) continue;
        uint2 ab = bd.m_ab[S.man_off + k];
        if ((L.flags[ab.x] & BF_TYPE) != 2u || (L.flags[ab.y] & BF_TYPE) != 2u) continue;
        uint32_t a = ab.x, b = ab.y;
        for (int guard = 0; guard < 100000; ++guard) {
            while (L.parent[a] != a) a = L.parent[a];
            while (L.parent[b] != b) b = L.parent[b];
            if (a == b) break;
            if (a > b) { uint32_t x = a; a = b; b = x; }
            // hook root b under a: 16-bit LDS CAS through the enclosing 32-bit word
            uint32_t* word = reinterpret_cast<uint32_t*>(&L.parent[b & ~1u]);
            uint32_t sh = (b & 1u) * 16u;
            uint32_t old = *word;
            if (((old >> sh) & 0xFFFFu) != b) continue;
            uint32_t nw = (old & ~(0xFFFFu << sh)) | (a << sh);
            if (atomicCAS(word, old, nw) == old) break;
        }
    }
    for (uint32_t jj = t; jj < S.n_j; jj += TB) {  // joints unite islands too (solver.h:820-825)
        const uint4 m = bd.j_meta[S.j_off + jj];
        if (!m.w || (L.flags[m.y] & BF_TYPE) != 2u || (L.flags[m.z] & BF_TYPE) != 2u) continue;
        uint32_t a = m.y, b = m.z;
        for (int guard = 0; guard < 100000; ++guard) {
            while (L.parent[a] != a) a = L.parent[a];
            while (L.parent[b] != b) b = L.parent[b];
            if (a == b) break;
            if (a > b) { uint32_t x = a; a = b; b = x; }
            uint32_t* word = reinterpret_cast<uint32_t*>(&L.parent[b & ~1u]);
            uint32_t sh = (b & 1u) * 16u;
            uint32_t old = *word;
            if (((old >> sh) & 0xFFFFu) != b) continue;
            uint32_t nw = (old & ~(0xFFFFu << sh)) | (a << sh);
            if (atomicCAS(word, old, nw) == old) break;
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < nb; i += TB) {
        uint32_t f = L.flags[i];
        if (!active_flag(f)) continue;
        v3 v = ld3(L.vel[i]), w = ld3(L.angv[i]);
        float4 pt = L.pos[i];
        if (dot(v, v) < cfg.sleep_lin * cfg.sleep_lin && dot(w, w) < cfg.sleep_ang * cfg.sleep_ang) pt.w = pt.w + cfg.h;
        else pt.w = 0.0f;
        L.pos[i] = pt;
    }
    __syncthreads();
    for (uint32_t k = t; k < np; k += TB) {
        if (L.colour[k] == kInvalid) continue;
        uint2 ab = bd.m_ab[S.man_off + k];
        for (int side = 0; side < 2; ++side) {
            uint32_t kin = side == 0 ? ab.x : ab.y, dyn = side == 0 ? ab.y : ab.x;
            uint32_t fk = L.flags[kin], fd = L.flags[dyn];
            if ((fk & BF_TYPE) != 1u || (fd & BF_TYPE) != 2u || (fd & BF_AWAKE)) continue;
            v3 kv = ld3(L.vel[kin]), kw = ld3(L.angv[kin]);
            if (dot(kv, kv) > cfg.sleep_lin * cfg.sleep_lin || dot(kw, kw) > cfg.sleep_ang * cfg.sleep_ang) {
                atomicOr(&L.flags[dyn], BF_AWAKE);
                L.pos[dyn].w = 0.0f;
            }
        }
    }
    for (uint32_t jj = t; jj < S.n_j; jj += TB) {  // and through joints (solver.h:854-858)
        const uint4 m = bd.j_meta[S.j_off + jj];
        if (!m.w) continue;
        for (int side = 0; side < 2; ++side) {
            uint32_t kin = side == 0 ? m.y : m.z, dyn = side == 0 ? m.z : m.y;
            uint32_t fk = L.flags[kin], fd = L.flags[dyn];
            if ((fk & BF_TYPE) != 1u || (fd & BF_TYPE) != 2u || (fd & BF_AWAKE)) continue;
            v3 kv = ld3(L.vel[kin]), kw = ld3(L.angv[kin]);
            if (dot(kv, kv) > cfg.sleep_lin * cfg.sleep_lin || dot(kw, kw) > cfg.sleep_ang * cfg.sleep_ang) {
                atomicOr(&L.flags[dyn], BF_AWAKE);
                L.pos[dyn].w = 0.0f;
            }
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < nb; i += TB) {
        if (!active_flag(L.flags[i])) continue;
        uint32_t r = i;
        while (L.parent[r] != r) r = L.parent[r];
        atomicMin(&L.isl_min[r], __float_as_uint(L.pos[i].w));
    }
    __syncthreads();
    uint32_t awake_n = 0;
    for (uint32_t i = t; i < nb; i += TB) {
        uint32_t f = L.flags[i];
        if (!active_flag(f)) continue;
        uint32_t r = i;
        while (L.parent[r] != r) r = L.parent[r];
        if (__uint_as_float(L.isl_min[r]) >= cfg.sleep_time) {
            L.flags[i] = f & ~BF_AWAKE;
            L.vel[i] = make_float4(0.0f, 0.0f, 0.0f, L.vel[i].w);
            L.angv[i] = make_float4(0.0f, 0.0f, 0.0f, L.angv[i].w);
        }
    }
    __syncthreads();
    // 13. prev pose, integrate velocities, world inertia (world.h:1303-1309, integrator.h:109-133)
    for (uint32_t i = t; i < nb; i += TB) {
        uint32_t f = L.flags[i];
        if (!(f & BF_ALIVE)) continue;
        float4 pos = L.pos[i], rot = L.rot[i];
        if (last) {
            bd.b_prev_pos[S.body_off + i] = make_float4(pos.x, pos.y, pos.z, 0.0f);
            bd.b_prev_rot[S.body_off + i] = rot;
        }
        bool dyn = (f & BF_TYPE) == 2u;
        if (dyn && (f & BF_AWAKE)) {
            ++awake_n;
            float4 vm = L.vel[i];
            v3 v = ld3(vm);
            float s2 = dot(v, v), m2 = cfg.max_lin_vel * cfg.max_lin_vel;
            if (s2 > m2 && s2 > kEps) {
                v = v * (cfg.max_lin_vel / sqrtf(s2));
                L.vel[i] = st3(v, vm.w);
            }
            L.pos[i] = st3(ld3(pos) + v * cfg.h, pos.w);
            quat q = ldq(rot);
            v3 w = ld3(L.angv[i]);
            if (cfg.exact_quat) {
                float angle = len(w) * cfg.h;
                if (angle > kEps) {
                    v3 ax = w / len(w);
                    float half = angle * 0.5f, sn = sinf(half);
                    quat dq;
                    dq.w = cosf(half); dq.x = ax.x * sn; dq.y = ax.y * sn; dq.z = ax.z * sn;
                    q = qnormalize(qmul(dq, q));
                }
            } else {
                quat wq;
                wq.w = 0.0f; wq.x = w.x; wq.y = w.y; wq.z = w.z;
                quat dq = qmul(wq, q);
                float hh = 0.5f * cfg.h;
                quat s;
                s.w = q.w + dq.w * hh; s.x = q.x + dq.x * hh; s.y = q.y + dq.y * hh; s.z = q.z + dq.z * hh;
                q = qnormalize(s);
            }
            L.rot[i] = stq(q);
            set_iiw(L, i, world_inertia(q, ld3(bd.b_iil[S.body_off + i])));
        } else if (!dyn) {
            m3 z;
            z.c0 = splat(0.0f); z.c1 = splat(0.0f); z.c2 = splat(0.0f);
            set_iiw(L, i, z);
        }
    }
    __syncthreads();
    if (S.n_j) {  // joint drift on the integrated poses, then world inertia (world.h:1310-1314)
        for (uint32_t it = 0; it < (cfg.pos_iters > 1u ? cfg.pos_iters : 1u); ++it) scene_joint_pass<TB>(bd, cfg, L, S, 2, false);
        for (uint32_t i = t; i < nb; i += TB) {
            uint32_t f = L.flags[i];
            if (!(f & BF_ALIVE)) continue;
            if ((f & BF_TYPE) == 2u) set_iiw(L, i, world_inertia(ldq(L.rot[i]), ld3(bd.b_iil[S.body_off + i])));
            else set_iiw(L, i, zero3());
        }
    }
    awake_n = wsum(awake_n);
    if (t == 0) L.awake = 0;
    __syncthreads();
    if ((t & 63) == 0) atomicAdd(&L.awake, awake_n);
    if (t == 0) { L.passes = passes; L.residual = resid; L.passes_total += passes; }
    __syncthreads();
}

template <int TB>
__global__ void __launch_bounds__(TB) k_batch(BatchDev bd, Cfg cfg) {
    const uint32_t* layer_mask = bd.layer_mask;
    const uint32_t gen0 = *bd.gen0;
    __shared__ Lds L;
    const uint32_t t = threadIdx.x;
    while (true) {
        __syncthreads();
        if (t == 0) {
            uint32_t q = atomicAdd(bd.queue, 1u);
            L.scene = q < bd.n_scenes ? bd.queue_order[q] : 0xFFFFFFFFu;
            L.err = 0;
            L.passes_total = 0;
        }
        __syncthreads();
        const uint32_t sc = L.scene;
        if (sc == 0xFFFFFFFFu) return;
        const Scene S = bd.scenes[sc];
        if (S.n_bodies > uint32_t(kMaxBodies) || S.n_shapes > uint32_t(kMaxBodies)) {
            if (t == 0) bd.stats[sc].err = SERR_SIZE;
            continue;
        }
        for (uint32_t i = t; i < S.n_bodies; i += TB) {
            uint32_t g = S.body_off + i;
            L.pos[i] = bd.b_pos[g];
            L.rot[i] = bd.b_rot[g];
            L.vel[i] = bd.b_vel[g];
            L.angv[i] = bd.b_angv[g];
            L.iiw[3 * i] = bd.b_iiw[3 * g];
            L.iiw[3 * i + 1] = bd.b_iiw[3 * g + 1];
            L.iiw[3 * i + 2] = bd.b_iiw[3 * g + 2];
            L.flags[i] = bd.b_flags[g];
        }
        for (uint32_t j = t; j < S.n_shapes; j += TB) {
            L.fmin[j] = bd.s_fmin[S.shape_off + j];
            L.fmax[j] = bd.s_fmax[S.shape_off + j];
        }
        __syncthreads();
        for (uint32_t k = 0; k < cfg.steps; ++k)
            scene_step<TB>(bd, cfg, L, S, gen0 + k, k + 1 == cfg.steps, layer_mask);
        for (uint32_t i = t; i < S.n_bodies; i += TB) {
            uint32_t g = S.body_off + i;
            bd.b_pos[g] = L.pos[i];
            bd.b_rot[g] = L.rot[i];
            bd.b_vel[g] = L.vel[i];
            bd.b_angv[g] = L.angv[i];
            bd.b_iiw[3 * g] = L.iiw[3 * i];
            bd.b_iiw[3 * g + 1] = L.iiw[3 * i + 1];
            bd.b_iiw[3 * g + 2] = L.iiw[3 * i + 2];
            bd.b_flags[g] = L.flags[i];
        }
        for (uint32_t j = t; j < S.n_shapes; j += TB) {
            bd.s_fmin[S.shape_off + j] = L.fmin[j];
            bd.s_fmax[S.shape_off + j] = L.fmax[j];
        }
        for (uint32_t c = t; c <= L.colours + 1 && c < uint32_t(kMaxColoursB + 2); c += TB) bd.m_colour_off[sc * (kMaxColoursB + 2) + c] = L.colour_off[c];
        for (uint32_t s = t; s < L.ordered; s += TB) bd.m_order[S.man_off + s] = L.order[s];
        if (t == 0) {
            SceneStats st;
            st.pairs = L.n_pairs; st.manifolds = L.n_valid; st.points = L.points; st.colours = L.colours;
            st.active = L.active; st.passes = L.passes; st.err = L.err; st.ordered = L.ordered;
            st.residual = L.residual; st.awake = L.awake;
            st.reserved[0] = bd.mlp ? (L.n_miss < kMissRows ? L.n_miss : kMissRows) : 0u;
            st.reserved[1] = bd.stats[sc].reserved[1] + L.passes_total;  // velocity passes since creation
            bd.stats[sc] = st;
        }
    }
}

}  // namespace pxb

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
namespace pxb {

__global__ void k_batch_prepare(BatchDev bd, uint32_t nb_total, uint32_t ns_total, float margin) {
    // world inverse inertia of every body (as add_body does, world.h:113) and each enabled shape's
    // proxy, fat = tight.expand(margin) (world.h:1369-1382)
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb_total) {
        uint32_t f = bd.b_flags[i];
        m3 m;
        if ((f & BF_TYPE) != 2u) { m.c0 = splat(0.0f); m.c1 = splat(0.0f); m.c2 = splat(0.0f); }
        else m = world_inertia(ldq(bd.b_rot[i]), ld3(bd.b_iil[i]));
        bd.b_iiw[3 * i] = make_float4(m.c0.x, m.c0.y, m.c0.z, m.c1.x);
        bd.b_iiw[3 * i + 1] = make_float4(m.c1.y, m.c1.z, m.c2.x, m.c2.y);
        bd.b_iiw[3 * i + 2] = make_float4(m.c2.z, 0.0f, 0.0f, 0.0f);
    }
}
__global__ void k_batch_proxies(BatchDev bd, uint32_t n_scenes, float margin) {
    uint32_t sc = blockIdx.x;
    if (sc >= n_scenes) return;
    Scene S = bd.scenes[sc];
    for (uint32_t j = threadIdx.x; j < S.n_shapes; j += blockDim.x) {
        uint4 info = bd.s_info[S.shape_off + j];
        uint32_t g = S.shape_off + j;
        if (!(info.y & SF_ENABLED) || info.x >= S.n_bodies) {
            bd.s_fmin[g] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            bd.s_fmax[g] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            continue;
        }
        uint32_t ob = S.body_off + info.x;
        ShapeGeom s = shape_geom(bd, g);
        v3 mn, mx;
        world_bounds(s, ld3(bd.b_pos[ob]), ldq(bd.b_rot[ob]), mn, mx);
        bd.s_fmin[g] = st3(mn - splat(margin), 0.0f);
        bd.s_fmax[g] = st3(mx + splat(margin), 0.0f);
    }
}
__global__ void k_batch_reset_queue(uint32_t* queue) { *queue = 0u; }
__global__ void k_batch_advance(uint32_t* gen, uint32_t steps) { *gen += steps; }


}  // namespace pxb

#include "pxg_hostio.h"

namespace {

thread_local std::string g_batch_error;
int bfail(int code, const std::string& m) {
    g_batch_error = "[pxg batch] " + m;
    return code;
}
#define BCHK(call)                                                                                        \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess) return bfail(PXG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

uint32_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return uint32_t(p);
}

}  // namespace

struct pxg_batch {
    pxg_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t n_scenes = 0, nb = 0, ns = 0;
    std::vector<pxb::Scene> scenes;
    std::vector<uint32_t> order_lpt, order_natural;
    pxb::BatchDev d{};
    uint32_t* gen = nullptr;
    uint32_t* layer = nullptr;
    uint32_t* ord_buf = nullptr;
    uint32_t host_gen = 2;
    std::vector<void*> allocs;
    int block = 256;
    bool longest_first = true;
    int blocks_per_cu[4] = {0, 0, 0, 0};  // TB = 64, 128, 256, 512
    int cus = 0;
    uint32_t wg_per_cu = 0;                // workgroups launched per CU (0: as many as fit)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t graph_steps = 0, graph_block = 0, graph_lpt = 2, graph_wpc = 0;
    float graph_h = 0.0f;
    pxg_config graph_cfg{};
    // joints and ignored pairs (scene-local body indices), body types for the joint layering
    std::vector<uint32_t> body_types;
    std::vector<pxg_joint> joints;
    std::vector<uint32_t> joint_offsets;
    std::vector<uint32_t> ignored;          // lo << 16 | hi
    std::vector<uint32_t> ignored_offsets;
    std::vector<void*> joint_allocs;
    pxg::MlpDev* mlp_dev = nullptr;   // the network's device image (learned warm start)
};

template <class T>
static int balloc(pxg_batch* b, T*& p, size_t count) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return bfail(PXG_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    b->allocs.push_back(q);
    p = static_cast<T*>(q);
    return PXG_OK;
}
#define BALLOC(ptr, n)                              \
    do {                                            \
        int r_ = balloc(b, (ptr), (n));             \
        if (r_ != PXG_OK) return r_;                \
    } while (0)
template <class T>
static int bupload(pxg_batch* b, T* dst, const std::vector<T>& src) {
    if (src.empty()) return PXG_OK;
    BCHK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, b->stream));
    return PXG_OK;
}

static void drop_graph(pxg_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}
static void batch_free(pxg_batch* b) {
    drop_graph(b);
    for (void* p : b->allocs) (void)hipFree(p);
    b->allocs.clear();
    for (void* p : b->joint_allocs) (void)hipFree(p);
    b->joint_allocs.clear();
}

// Per scene: live joints, their forward and reverse layers, and the sorted body pairs without
// contacts (jointed without collide_connected, or ignored); uploads them with the scene table.
static int rebuild_batch_joints(pxg_batch* b) {
    (void)hipStreamSynchronize(b->stream);
    drop_graph(b);  // the captured launch holds the old buffers
    for (void* p : b->joint_allocs) (void)hipFree(p);
    b->joint_allocs.clear();
    const uint32_t nj = uint32_t(b->joints.size());
    std::vector<uint4> meta(nj);
    std::vector<float4> def(6 * size_t(nj)), acc(2 * size_t(nj));
    std::vector<uint16_t> order(2 * size_t(nj));
    std::vector<uint32_t> loff, keys;
    for (uint32_t k = 0; k < b->n_scenes; ++k) {
        pxb::Scene& S = b->scenes[k];
        const uint32_t j0 = b->joint_offsets[k], n = b->joint_offsets[k + 1] - j0;
        std::vector<uint32_t> types(b->body_types.begin() + S.body_off, b->body_types.begin() + S.body_off + S.n_bodies);
        std::vector<uint8_t> live(n, 0);
        std::vector<uint32_t> sk(b->ignored.begin() + b->ignored_offsets[k], b->ignored.begin() + b->ignored_offsets[k + 1]);
        for (uint32_t q = 0; q < n; ++q) {
            const pxg_joint& jt = b->joints[j0 + q];
            bool in_range = jt.body_a < S.n_bodies && jt.body_b < S.n_bodies;
            live[q] = (jt.valid && jt.enabled && in_range) ? 1 : 0;
            meta[j0 + q] = make_uint4(jt.type, in_range ? jt.body_a : 0u, in_range ? jt.body_b : 0u, live[q]);
            pxg::pack_joint(jt, &def[6 * size_t(j0 + q)], &acc[2 * size_t(j0 + q)]);
            if (jt.valid && !jt.collide_connected && in_range)
                sk.push_back((std::min(jt.body_a, jt.body_b) << 16) | std::max(jt.body_a, jt.body_b));
        }
        std::vector<uint32_t> of, orv, lf, lr;
        uint32_t Lf = pxg::layer_joint_list(b->joints.data() + j0, n, live, types, false, of, lf);
        uint32_t Lr = pxg::layer_joint_list(b->joints.data() + j0, n, live, types, true, orv, lr);
        uint32_t L = std::max(Lf, Lr);
        lf.resize(L + 1, lf.back());
        lr.resize(L + 1, lr.back());
        for (uint32_t q = 0; q < n; ++q) {
            order[2 * size_t(j0) + q] = uint16_t(of[q]);
            order[2 * size_t(j0) + n + q] = uint16_t(orv[q]);
        }
        S.j_off = j0;
        S.n_j = n;
        S.j_layers = L;
        S.jl_off = uint32_t(loff.size());
        loff.insert(loff.end(), lf.begin(), lf.end());
        loff.insert(loff.end(), lr.begin(), lr.end());
        std::sort(sk.begin(), sk.end());
        sk.erase(std::unique(sk.begin(), sk.end()), sk.end());
        S.nc_off = uint32_t(keys.size());
        S.n_nc = uint32_t(sk.size());
        keys.insert(keys.end(), sk.begin(), sk.end());
    }
    auto up = [&](auto*& dst, const auto& src) -> int {
        using T = typename std::remove_reference<decltype(src)>::type::value_type;
        void* q = nullptr;
        size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
        if (hipMalloc(&q, bytes) != hipSuccess) return bfail(PXG_ERR_HIP, "hipMalloc (joints)");
        b->joint_allocs.push_back(q);
        if (!src.empty() && hipMemcpyAsync(q, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, b->stream) != hipSuccess)
            return bfail(PXG_ERR_HIP, "hipMemcpyAsync (joints)");
        dst = static_cast<typename std::remove_reference<decltype(dst)>::type>(q);
        return PXG_OK;
    };
    pxb::BatchDev& d = b->d;
    const uint4* jm; const float4* jd; float4* ja; const uint16_t* jo; const uint32_t* jl; const uint32_t* nk;
    int r;
    if ((r = up(jm, meta)) || (r = up(jd, def)) || (r = up(ja, acc)) || (r = up(jo, order)) || (r = up(jl, loff)) ||
        (r = up(nk, keys)))
        return r;
    d.j_meta = jm; d.j_def = jd; d.j_acc = ja; d.j_order = jo; d.j_loff = jl; d.nc_keys = nk;
    BCHK(hipMemcpyAsync(const_cast<pxb::Scene*>(d.scenes), b->scenes.data(), b->scenes.size() * sizeof(pxb::Scene),
                        hipMemcpyHostToDevice, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    return PXG_OK;
}

namespace pxg { void mlp_to_device_image(const pxg_mlp* net, MlpDev* img); }

// The learned warm start for every scene (the single world's pxg_set_mlp); net == nullptr turns it off.
extern "C" int pxg_batch_set_mlp(pxg_batch* b, const pxg_mlp* net) {
    if (!b) return bfail(PXG_ERR_INVALID, "pxg_batch_set_mlp: null handle");
    (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (!net) {
        b->d.mlp = nullptr;
        return PXG_OK;
    }
    if (!b->d.feat) {
        const size_t rows = size_t(b->n_scenes) * pxb::kMissRows;
        pxg::MlpDev* img_dev = nullptr;
        BALLOC(img_dev, 1);
        BALLOC(b->d.feat, rows * pxg::kMlpIn);
        BALLOC(b->d.miss_pt, rows);
        BALLOC(b->d.pred, rows * pxg::kMlpOut);
        b->mlp_dev = img_dev;
    }
    pxg::MlpDev img;
    pxg::mlp_to_device_image(net, &img);
    BCHK(hipMemcpyAsync(b->mlp_dev, &img, sizeof(img), hipMemcpyHostToDevice, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    b->d.mlp = b->mlp_dev;
    return PXG_OK;
}

// The last step's predictions of one scene, sorted by warm-start key (pack_warm_start_key,
// dynamics/solver.h:46-51, scene-local body indices): the clamped initial impulses of its misses.
extern "C" int pxg_batch_get_predictions(pxg_batch* b, uint32_t scene, uint64_t* keys, float* imp3, uint32_t cap,
                                         uint32_t* n) {
    if (!b || !n || scene >= b->n_scenes) return bfail(PXG_ERR_INVALID, "pxg_batch_get_predictions: bad argument");
    *n = 0;
    if (!b->d.mlp) return PXG_OK;
    hipStream_t st = b->stream;
    pxb::SceneStats ss;
    BCHK(hipMemcpyAsync(&ss, b->d.stats + scene, sizeof(ss), hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    const uint32_t P = ss.reserved[0];
    if (!P) return PXG_OK;
    const size_t base = size_t(scene) * pxb::kMissRows;
    std::vector<uint32_t> pt(P);
    std::vector<float> pr(3 * size_t(P));
    BCHK(hipMemcpyAsync(pt.data(), b->d.miss_pt + base, P * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(pr.data(), b->d.pred + base * 3, pr.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    std::vector<std::pair<uint64_t, uint32_t>> e;
    for (uint32_t j = 0; j < P; ++j) {
        const uint32_t gm = pt[j] / 4u, q = pt[j] % 4u;
        uint2 ab;
        float4 qq;
        BCHK(hipMemcpy(&ab, b->d.m_ab + gm, sizeof(ab), hipMemcpyDeviceToHost));
        BCHK(hipMemcpy(&qq, b->d.m_q + 4 * size_t(gm) + q, sizeof(qq), hipMemcpyDeviceToHost));
        uint32_t fid;
        std::memcpy(&fid, &qq.y, 4);
        const uint32_t lo = std::min(ab.x, ab.y), hi = std::max(ab.x, ab.y);
        e.push_back({(uint64_t(lo) << 40) ^ (uint64_t(hi) << 12) ^ uint64_t(fid), j});
    }
    std::sort(e.begin(), e.end());
    *n = P;
    for (uint32_t i = 0; i < P && i < cap; ++i) {
        if (keys) keys[i] = e[i].first;
        if (imp3) for (int k = 0; k < 3; ++k) imp3[3 * i + k] = pr[3 * size_t(e[i].second) + k];
    }
    return PXG_OK;
}

extern "C" int pxg_batch_set_joints(pxg_batch* b, const uint32_t* joint_offsets, const pxg_joint* joints) {
    if (!b || !joint_offsets || (!joints && joint_offsets[b->n_scenes]))
        return bfail(PXG_ERR_INVALID, "pxg_batch_set_joints: null argument");
    for (uint32_t k = 0; k < b->n_scenes; ++k) {
        if (joint_offsets[k + 1] < joint_offsets[k]) return bfail(PXG_ERR_INVALID, "pxg_batch_set_joints: offsets must be non-decreasing");
        if (joint_offsets[k + 1] - joint_offsets[k] > 4096u)
            return bfail(PXG_ERR_CAPACITY, "pxg_batch_set_joints: more than 4096 joints in scene " + std::to_string(k));
        for (uint32_t j = joint_offsets[k]; j < joint_offsets[k + 1]; ++j) {
            const pxg_joint& jt = joints[j];
            if (jt.type > PXG_CONE_TWIST) return bfail(PXG_ERR_INVALID, "pxg_batch_set_joints: bad joint type");
            if (jt.body_a >= b->scenes[k].n_bodies || jt.body_b >= b->scenes[k].n_bodies)
                return bfail(PXG_ERR_INVALID, "pxg_batch_set_joints: body index beyond scene " + std::to_string(k));
            if (jt.body_a == jt.body_b) return bfail(PXG_ERR_INVALID, "pxg_batch_set_joints: a joint needs two different bodies");
        }
    }
    b->joint_offsets.assign(joint_offsets, joint_offsets + b->n_scenes + 1);
    b->joints.assign(joints, joints + joint_offsets[b->n_scenes]);
    return rebuild_batch_joints(b);
}

extern "C" int pxg_batch_get_joints(pxg_batch* b, pxg_joint* out, uint32_t n) {
    if (!b || (!out && n)) return bfail(PXG_ERR_INVALID, "pxg_batch_get_joints: null argument");
    if (n > b->joints.size()) return bfail(PXG_ERR_INVALID, "pxg_batch_get_joints: more joints requested than the batch holds");
    if (!n) return PXG_OK;
    std::vector<float4> acc(2 * size_t(n));
    BCHK(hipMemcpyAsync(acc.data(), b->d.j_acc, acc.size() * sizeof(float4), hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    for (uint32_t j = 0; j < n; ++j) {
        pxg::unpack_joint_acc(&acc[2 * size_t(j)], b->joints[j]);
        out[j] = b->joints[j];
    }
    return PXG_OK;
}

extern "C" int pxg_batch_set_ignored_pairs(pxg_batch* b, const uint32_t* pair_offsets, const uint32_t* ab) {
    if (!b || !pair_offsets || (!ab && pair_offsets[b->n_scenes]))
        return bfail(PXG_ERR_INVALID, "pxg_batch_set_ignored_pairs: null argument");
    b->ignored.clear();
    b->ignored_offsets.assign(b->n_scenes + 1, 0u);
    for (uint32_t k = 0; k < b->n_scenes; ++k) {
        if (pair_offsets[k + 1] < pair_offsets[k]) return bfail(PXG_ERR_INVALID, "pxg_batch_set_ignored_pairs: offsets must be non-decreasing");
        for (uint32_t i = pair_offsets[k]; i < pair_offsets[k + 1]; ++i) {
            uint32_t x = ab[2 * i], y = ab[2 * i + 1];
            if (x >= b->scenes[k].n_bodies || y >= b->scenes[k].n_bodies)
                return bfail(PXG_ERR_INVALID, "pxg_batch_set_ignored_pairs: body index beyond scene " + std::to_string(k));
            if (x != y) b->ignored.push_back((std::min(x, y) << 16) | std::max(x, y));
        }
        b->ignored_offsets[k + 1] = uint32_t(b->ignored.size());
    }
    return rebuild_batch_joints(b);
}

extern "C" const char* pxg_batch_last_error(void) { return g_batch_error.c_str(); }

extern "C" int pxg_batch_create(const pxg_config* cfg, uint32_t n_scenes, const uint32_t* body_offsets,
                                const uint32_t* shape_offsets, const pxg_body* bodies, const pxg_shape* shapes,
                                pxg_batch** out) {
    if (!out || !body_offsets || !shape_offsets || (!bodies && body_offsets[n_scenes]) ||
        (!shapes && shape_offsets[n_scenes]))
        return bfail(PXG_ERR_INVALID, "pxg_batch_create: null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return bfail(PXG_ERR_HIP, "pxg_batch_create: no HIP device");
    pxg_batch* b = new pxg_batch();
    if (cfg) b->cfg = *cfg;
    else pxg_config_default(&b->cfg);
    (void)hipGetDevice(&b->device);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return bfail(PXG_ERR_HIP, "pxg_batch_create: hipStreamCreate failed");
    }
    auto bail = [&](int code) {
        batch_free(b);
        (void)hipStreamDestroy(b->stream);
        delete b;
        return code;
    };
    b->n_scenes = n_scenes;
    b->nb = body_offsets[n_scenes];
    b->ns = shape_offsets[n_scenes];
    // scene regions
    b->scenes.resize(n_scenes);
    uint64_t man = 0, cache = 0;
    for (uint32_t k = 0; k < n_scenes; ++k) {
        pxb::Scene& S = b->scenes[k];
        S.body_off = body_offsets[k];
        S.n_bodies = body_offsets[k + 1] - body_offsets[k];
        S.shape_off = shape_offsets[k];
        S.n_shapes = shape_offsets[k + 1] - shape_offsets[k];
        if (body_offsets[k + 1] < body_offsets[k] || shape_offsets[k + 1] < shape_offsets[k])
            return bail(bfail(PXG_ERR_INVALID, "pxg_batch_create: offsets must be non-decreasing"));
        if (S.n_bodies > uint32_t(pxb::kMaxBodies) || S.n_shapes > uint32_t(pxb::kMaxBodies))
            return bail(bfail(PXG_ERR_CAPACITY, "pxg_batch_create: scene " + std::to_string(k) + " has more than " +
                                                    std::to_string(pxb::kMaxBodies) + " bodies or shapes"));
        S.man_off = uint32_t(man);
        S.man_cap = std::min<uint32_t>(uint32_t(pxb::kMaxPairs), 12u * S.n_shapes + 64u);
        man += S.man_cap;
        uint32_t tc = next_pow2(16ull * S.n_shapes + 64ull);
        S.cache_off = uint32_t(cache);
        S.cache_mask = tc - 1u;
        cache += tc;
    }
    if (man >= (1ull << 31) || cache >= (1ull << 31)) return bail(bfail(PXG_ERR_CAPACITY, "pxg_batch_create: batch too large"));
    // LPT queue: scenes by decreasing body count
    b->order_natural.resize(n_scenes);
    for (uint32_t k = 0; k < n_scenes; ++k) b->order_natural[k] = k;
    b->order_lpt = b->order_natural;
    std::stable_sort(b->order_lpt.begin(), b->order_lpt.end(),
                     [&](uint32_t x, uint32_t y) { return b->scenes[x].n_bodies > b->scenes[y].n_bodies; });
    // bodies and shapes
    pxg::BodySoA bs;
    if (!pxg::pack_bodies(bodies, b->nb, bs)) return bail(bfail(PXG_ERR_INVALID, "pxg_batch_create: bad body type"));
    b->body_types.resize(b->nb);
    for (uint32_t i = 0; i < b->nb; ++i) b->body_types[i] = bodies[i].type;
    b->joint_offsets.assign(n_scenes + 1, 0u);
    b->ignored_offsets.assign(n_scenes + 1, 0u);
    pxg::ShapeSoA ss;
    std::vector<uint32_t> shape_scene(b->ns);
    for (uint32_t k = 0; k < n_scenes; ++k)
        for (uint32_t j = shape_offsets[k]; j < shape_offsets[k + 1]; ++j) shape_scene[j] = k;
    for (uint32_t j = 0; j < b->ns; ++j)
        if (shapes[j].owner >= body_offsets[shape_scene[j] + 1] - body_offsets[shape_scene[j]])
            return bail(bfail(PXG_ERR_INVALID, "pxg_batch_create: shape owner outside its scene"));
    const char* e = pxg::pack_shapes(shapes, b->ns, 0xFFFFFFFFu, [&](size_t j) {
        return bodies[body_offsets[shape_scene[j]] + shapes[j].owner].type == PXG_STATIC;
    }, ss);
    if (e) return bail(bfail(PXG_ERR_INVALID, std::string("pxg_batch_create: ") + e));
    pxb::BatchDev& d = b->d;
    d.n_scenes = n_scenes;
    pxb::Scene* dsc;
    uint32_t *dq, *dqo;
    pxb::SceneStats* dst;
    int r;
#define BA(ptr, n) do { if ((r = balloc(b, (ptr), (n))) != PXG_OK) return bail(r); } while (0)
    BA(dsc, n_scenes); BA(dqo, 2 * size_t(n_scenes)); BA(dq, 1); BA(dst, n_scenes);
    d.scenes = dsc; d.queue_order = dqo; d.queue = dq; d.stats = dst;
    size_t nbt = b->nb, nst = b->ns;
    BA(d.b_pos, nbt); BA(d.b_rot, nbt); BA(d.b_vel, nbt); BA(d.b_angv, nbt); BA(d.b_force, nbt); BA(d.b_torque, nbt);
    BA(d.b_iil, nbt); BA(d.b_iiw, 3 * nbt); BA(d.b_prev_pos, nbt); BA(d.b_prev_rot, nbt); BA(d.b_flags, nbt);
    uint4* sinfo; float4 *slc, *slr, *sdim, *smat;
    BA(sinfo, nst); BA(slc, nst); BA(slr, nst); BA(sdim, nst); BA(smat, nst);
    d.s_info = sinfo; d.s_lc = slc; d.s_lr = slr; d.s_dim = sdim; d.s_mat = smat;
    BA(d.s_fmin, nst); BA(d.s_fmax, nst);
    BA(d.m_valid, man); BA(d.m_pair, man); BA(d.m_ab, man); BA(d.m_nf, man); BA(d.m_meta, man);
    BA(d.m_p, 4 * man); BA(d.m_q, 4 * man); BA(d.m_imp, 4 * man); BA(d.m_order, man);
    BA(d.m_colour_off, size_t(n_scenes) * (pxb::kMaxColoursB + 2));
    BA(d.c_key, cache); BA(d.c_ord, cache); BA(d.c_val, cache);
    BA(b->gen, 1); BA(b->layer, 32);
    d.gen0 = b->gen;
    d.layer_mask = b->layer;
#undef BA
    hipStream_t st = b->stream;
    std::vector<uint32_t> qo(2 * size_t(n_scenes));
    std::copy(b->order_lpt.begin(), b->order_lpt.end(), qo.begin());
    std::copy(b->order_natural.begin(), b->order_natural.end(), qo.begin() + n_scenes);
    std::vector<uint32_t> masks(32, 0xFFFFFFFFu);
    std::vector<uint32_t> gen0(1, b->host_gen);
    if ((r = bupload(b, dsc, b->scenes)) || (r = bupload(b, dqo, qo)) || (r = bupload(b, d.b_pos, bs.pos)) ||
        (r = bupload(b, d.b_rot, bs.rot)) || (r = bupload(b, d.b_vel, bs.vel)) || (r = bupload(b, d.b_angv, bs.angv)) ||
        (r = bupload(b, d.b_force, bs.force)) || (r = bupload(b, d.b_torque, bs.torque)) ||
        (r = bupload(b, d.b_iil, bs.iil)) || (r = bupload(b, d.b_prev_pos, bs.ppos)) ||
        (r = bupload(b, d.b_prev_rot, bs.prot)) || (r = bupload(b, d.b_flags, bs.flags)) ||
        (r = bupload(b, sinfo, ss.info)) || (r = bupload(b, slc, ss.lc)) || (r = bupload(b, slr, ss.lr)) ||
        (r = bupload(b, sdim, ss.dim)) || (r = bupload(b, smat, ss.mat)) || (r = bupload(b, b->layer, masks)) ||
        (r = bupload(b, b->gen, gen0)))
        return bail(r);
    if (hipMemsetAsync(d.c_key, 0, cache * sizeof(unsigned long long), st) != hipSuccess ||
        hipMemsetAsync(d.c_ord, 0, cache * sizeof(unsigned long long), st) != hipSuccess ||
        hipMemsetAsync(dst, 0, n_scenes * sizeof(pxb::SceneStats), st) != hipSuccess)
        return bail(bfail(PXG_ERR_HIP, "pxg_batch_create: hipMemsetAsync failed"));
    if (b->nb) hipLaunchKernelGGL(pxb::k_batch_prepare, dim3((b->nb + 255) / 256), dim3(256), 0, st, d, b->nb, b->ns, b->cfg.aabb_margin);
    if (n_scenes) hipLaunchKernelGGL(pxb::k_batch_proxies, dim3(n_scenes), dim3(256), 0, st, d, n_scenes, b->cfg.aabb_margin);
    if (hipStreamSynchronize(st) != hipSuccess) return bail(bfail(PXG_ERR_HIP, "pxg_batch_create: setup kernels failed"));
    if ((r = rebuild_batch_joints(b)) != PXG_OK) return bail(r);  // empty joint table
    (void)hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->blocks_per_cu[0], reinterpret_cast<const void*>(&pxb::k_batch<64>), 64, 0);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->blocks_per_cu[1], reinterpret_cast<const void*>(&pxb::k_batch<128>), 128, 0);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->blocks_per_cu[2], reinterpret_cast<const void*>(&pxb::k_batch<256>), 256, 0);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->blocks_per_cu[3], reinterpret_cast<const void*>(&pxb::k_batch<512>), 512, 0);
    *out = b;
    return PXG_OK;
}

extern "C" int pxg_batch_destroy(pxg_batch* b) {
    if (!b) return bfail(PXG_ERR_INVALID, "pxg_batch_destroy: null handle");
    (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return PXG_OK;
}

extern "C" int pxg_batch_set_tuning(pxg_batch* b, uint32_t block_size, uint32_t longest_first) {
    if (!b) return bfail(PXG_ERR_INVALID, "pxg_batch_set_tuning: null handle");
    if (block_size != 64 && block_size != 128 && block_size != 256 && block_size != 512)
        return bfail(PXG_ERR_INVALID, "pxg_batch_set_tuning: block size must be 64, 128, 256 or 512");
    b->block = int(block_size);
    b->longest_first = longest_first != 0;
    return PXG_OK;
}

extern "C" int pxg_batch_set_occupancy(pxg_batch* b, uint32_t workgroups_per_cu) {
    if (!b) return bfail(PXG_ERR_INVALID, "pxg_batch_set_occupancy: null handle");
    b->wg_per_cu = workgroups_per_cu;
    return PXG_OK;
}

extern "C" int pxg_batch_set_layer_masks(pxg_batch* b, const uint32_t masks[32]) {
    if (!b || !masks) return bfail(PXG_ERR_INVALID, "pxg_batch_set_layer_masks: null argument");
    BCHK(hipMemcpyAsync(b->layer, masks, 32 * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    return PXG_OK;
}

template <int TB>
static hipError_t batch_launch(pxg_batch* b, const pxb::Cfg& c, int bpc) {
    pxb::BatchDev d = b->d;
    if (!b->longest_first) d.queue_order = b->d.queue_order + b->n_scenes;
    const uint32_t fit = uint32_t(std::max(1, bpc));
    const uint32_t per_cu = b->wg_per_cu ? std::min(b->wg_per_cu, fit) : fit;
    uint32_t grid = std::max(1u, std::min<uint32_t>(b->n_scenes, per_cu * uint32_t(std::max(1, b->cus))));
    hipLaunchKernelGGL(pxb::k_batch<TB>, dim3(grid), dim3(TB), 0, b->stream, d, c);
    return hipGetLastError();
}

// Steps every scene `steps` times in one persistent launch, replayed from a captured HIP graph
// (queue reset, the batch kernel, the generation advance).
extern "C" int pxg_batch_step(pxg_batch* b, float h, uint32_t steps) {
    if (!b) return bfail(PXG_ERR_INVALID, "pxg_batch_step: null handle");
    if (!(h > 0.0f) || steps == 0) return bfail(PXG_ERR_INVALID, "pxg_batch_step: h > 0 and steps > 0 required");
    if (b->host_gen + steps >= 0x4000u) {
        // the cache keys carry 14 generation bits: restart the generation count with empty tables
        (void)hipStreamSynchronize(b->stream);
        uint64_t cache = 0;
        for (auto& S : b->scenes) cache += S.cache_mask + 1ull;
        BCHK(hipMemsetAsync(b->d.c_key, 0, cache * sizeof(unsigned long long), b->stream));
        BCHK(hipMemsetAsync(b->d.c_ord, 0, cache * sizeof(unsigned long long), b->stream));
        b->host_gen = 2;
        BCHK(hipMemcpyAsync(b->gen, &b->host_gen, 4, hipMemcpyHostToDevice, b->stream));
        BCHK(hipStreamSynchronize(b->stream));
    }
    const pxg_config& g = b->cfg;
    bool rebuild = !b->exec || b->graph_steps != steps || b->graph_block != uint32_t(b->block) ||
                   b->graph_lpt != uint32_t(b->longest_first) || b->graph_wpc != b->wg_per_cu || b->graph_h != h ||
                   std::memcmp(&b->graph_cfg, &g, sizeof(g)) != 0;
    if (rebuild) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        pxb::Cfg c;
        c.h = h; c.gx = g.gravity[0]; c.gy = g.gravity[1]; c.gz = g.gravity[2]; c.margin = g.aabb_margin;
        c.pos_corr = g.position_correction; c.slop = g.linear_slop; c.max_corr = g.max_linear_correction;
        c.rest_thresh = g.restitution_threshold; c.sleep_lin = g.sleep_linear; c.sleep_ang = g.sleep_angular;
        c.sleep_time = g.sleep_time; c.max_lin_vel = g.max_linear_velocity; c.exact_quat = g.use_exact_quaternion_integration;
        c.vel_iters = g.velocity_iterations; c.relax_iters = g.relax_iterations; c.pos_iters = g.position_iterations;
        c.residual_mode = g.residual_mode; c.residual_max = g.residual_max_iterations ? g.residual_max_iterations : 256u;
        c.residual_tol = g.residual_tolerance; c.steps = steps;
        c.check_every = g.residual_check_interval ? g.residual_check_interval : 1u;
        BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipLaunchKernelGGL(pxb::k_batch_reset_queue, dim3(1), dim3(1), 0, b->stream, b->d.queue);
        hipError_t le = hipSuccess;
        switch (b->block) {
            case 64: le = batch_launch<64>(b, c, b->blocks_per_cu[0]); break;
            case 128: le = batch_launch<128>(b, c, b->blocks_per_cu[1]); break;
            case 512: le = batch_launch<512>(b, c, b->blocks_per_cu[3]); break;
            default: le = batch_launch<256>(b, c, b->blocks_per_cu[2]); break;
        }
        hipLaunchKernelGGL(pxb::k_batch_advance, dim3(1), dim3(1), 0, b->stream, b->gen, steps);
        hipGraph_t graph = nullptr;
        hipError_t ce = hipStreamEndCapture(b->stream, &graph);
        if (le != hipSuccess || ce != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return bfail(PXG_ERR_HIP, std::string("pxg_batch_step: capture failed: ") + hipGetErrorString(le != hipSuccess ? le : ce));
        }
        b->graph = graph;
        BCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_steps = steps;
        b->graph_block = uint32_t(b->block);
        b->graph_lpt = uint32_t(b->longest_first);
        b->graph_wpc = b->wg_per_cu;
        b->graph_h = h;
        b->graph_cfg = g;
    }
    BCHK(hipGraphLaunch(b->exec, b->stream));
    b->host_gen += steps;
    return PXG_OK;
}

extern "C" int pxg_batch_synchronize(pxg_batch* b) {
    if (!b) return bfail(PXG_ERR_INVALID, "pxg_batch_synchronize: null handle");
    BCHK(hipStreamSynchronize(b->stream));
    return PXG_OK;
}

extern "C" int pxg_batch_get_bodies(pxg_batch* b, pxg_body* out, uint32_t n) {
    if (!b || (!out && n)) return bfail(PXG_ERR_INVALID, "pxg_batch_get_bodies: null argument");
    if (n > b->nb) return bfail(PXG_ERR_INVALID, "pxg_batch_get_bodies: n exceeds the body count");
    if (!n) return PXG_OK;
    pxg::BodySoA s;
    s.resize(n);
    std::vector<float4> iiw(3 * size_t(n));
    hipStream_t st = b->stream;
    const pxb::BatchDev& d = b->d;
    BCHK(hipMemcpyAsync(s.pos.data(), d.b_pos, n * sizeof(float4), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(s.rot.data(), d.b_rot, n * sizeof(float4), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(s.vel.data(), d.b_vel, n * sizeof(float4), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(s.angv.data(), d.b_angv, n * sizeof(float4), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(s.force.data(), d.b_force, n * sizeof(float4), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(s.torque.data(), d.b_torque, n * sizeof(float4), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(s.iil.data(), d.b_iil, n * sizeof(float4), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(iiw.data(), d.b_iiw, 3 * n * sizeof(float4), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(s.ppos.data(), d.b_prev_pos, n * sizeof(float4), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(s.prot.data(), d.b_prev_rot, n * sizeof(float4), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(s.flags.data(), d.b_flags, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    pxg::unpack_bodies(s, iiw, n, out);
    return PXG_OK;
}

extern "C" int pxg_batch_get_stats(pxg_batch* b, pxg_batch_scene_stats* out, uint32_t n) {
    if (!b || (!out && n)) return bfail(PXG_ERR_INVALID, "pxg_batch_get_stats: null argument");
    if (n > b->n_scenes) return bfail(PXG_ERR_INVALID, "pxg_batch_get_stats: n exceeds the scene count");
    static_assert(sizeof(pxg_batch_scene_stats) == sizeof(pxb::SceneStats), "stats layout");
    BCHK(hipMemcpyAsync(out, b->d.stats, n * sizeof(pxb::SceneStats), hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    return PXG_OK;
}

// Valid manifolds of one scene's last step, in pair order (the reference's manifolds() order),
// with scene-local body and shape indices and the solver's final impulses.
extern "C" int pxg_batch_get_manifolds(pxg_batch* b, uint32_t scene, pxg_manifold* out, uint32_t cap, uint32_t* n) {
    if (!b || !n || scene >= b->n_scenes) return bfail(PXG_ERR_INVALID, "pxg_batch_get_manifolds: bad argument");
    const pxb::Scene& S = b->scenes[scene];
    pxb::SceneStats stt;
    hipStream_t st = b->stream;
    BCHK(hipMemcpyAsync(&stt, b->d.stats + scene, sizeof(stt), hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    uint32_t P = std::min(stt.pairs, S.man_cap);
    std::vector<uint32_t> valid(P);
    std::vector<uint2> pr(P), ab(P);
    std::vector<float4> nf(P), p(4 * size_t(P)), q(4 * size_t(P)), imp(4 * size_t(P));
    std::vector<uint4> meta(P);
    if (P) {
        const pxb::BatchDev& d = b->d;
        size_t o = S.man_off;
        BCHK(hipMemcpyAsync(valid.data(), d.m_valid + o, P * 4, hipMemcpyDeviceToHost, st));
        BCHK(hipMemcpyAsync(pr.data(), d.m_pair + o, P * sizeof(uint2), hipMemcpyDeviceToHost, st));
        BCHK(hipMemcpyAsync(ab.data(), d.m_ab + o, P * sizeof(uint2), hipMemcpyDeviceToHost, st));
        BCHK(hipMemcpyAsync(nf.data(), d.m_nf + o, P * sizeof(float4), hipMemcpyDeviceToHost, st));
        BCHK(hipMemcpyAsync(meta.data(), d.m_meta + o, P * sizeof(uint4), hipMemcpyDeviceToHost, st));
        BCHK(hipMemcpyAsync(p.data(), d.m_p + 4 * o, 4 * P * sizeof(float4), hipMemcpyDeviceToHost, st));
        BCHK(hipMemcpyAsync(q.data(), d.m_q + 4 * o, 4 * P * sizeof(float4), hipMemcpyDeviceToHost, st));
        BCHK(hipMemcpyAsync(imp.data(), d.m_imp + 4 * o, 4 * P * sizeof(float4), hipMemcpyDeviceToHost, st));
        BCHK(hipStreamSynchronize(st));
    }
    uint32_t m = 0;
    for (uint32_t k = 0; k < P; ++k) {
        if (!valid[k]) continue;
        if (out && m < cap) {
            pxg_manifold& o = out[m];
            std::memset(&o, 0, sizeof(o));
            o.a = ab[k].x; o.b = ab[k].y; o.shape_a = pr[k].x; o.shape_b = pr[k].y;
            o.normal[0] = nf[k].x; o.normal[1] = nf[k].y; o.normal[2] = nf[k].z; o.friction = nf[k].w;
            std::memcpy(&o.restitution, &meta[k].y, 4);
            o.count = meta[k].x & 255u;
            o.is_trigger = (meta[k].x >> 8) & 1u;
            o.valid = 1;
            for (int t = 0; t < 4; ++t) {
                pxg_contact& c = o.points[t];
                const float4 &pp = p[4 * k + t], &qq = q[4 * k + t], &im = imp[4 * k + t];
                c.position[0] = pp.x; c.position[1] = pp.y; c.position[2] = pp.z; c.penetration = pp.w;
                c.separation = qq.x;
                std::memcpy(&c.feature_id, &qq.y, 4);
                c.restitution_bias = qq.z;
                if (t < int(o.count) && !o.is_trigger) {
                    c.normal_impulse = im.x; c.tangent_impulse[0] = im.y; c.tangent_impulse[1] = im.z;
                }
            }
        }
        ++m;
    }
    *n = m;
    return PXG_OK;
}

// One scene's last solve order as indices into pxg_batch_get_manifolds' list, colour by colour,
// then the manifolds the solver skipped.
extern "C" int pxg_batch_get_solve_order(pxg_batch* b, uint32_t scene, uint32_t* order, uint32_t cap, uint32_t* n,
                                         uint32_t* colour_offsets, uint32_t colour_cap, uint32_t* n_colours) {
    if (!b || !n || !n_colours || scene >= b->n_scenes) return bfail(PXG_ERR_INVALID, "pxg_batch_get_solve_order: bad argument");
    const pxb::Scene& S = b->scenes[scene];
    pxb::SceneStats stt;
    hipStream_t st = b->stream;
    BCHK(hipMemcpyAsync(&stt, b->d.stats + scene, sizeof(stt), hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    uint32_t P = std::min(stt.pairs, S.man_cap), O = std::min(stt.ordered, S.man_cap);
    std::vector<uint32_t> valid(P), offs(pxb::kMaxColoursB + 2);
    std::vector<uint16_t> ord(O);
    if (P) BCHK(hipMemcpyAsync(valid.data(), b->d.m_valid + S.man_off, P * 4, hipMemcpyDeviceToHost, st));
    if (O) BCHK(hipMemcpyAsync(ord.data(), b->d.m_order + S.man_off, O * 2, hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(offs.data(), b->d.m_colour_off + size_t(scene) * (pxb::kMaxColoursB + 2),
                        offs.size() * 4, hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    std::vector<uint32_t> rank(P, 0xFFFFFFFFu);
    uint32_t m = 0;
    for (uint32_t k = 0; k < P; ++k)
        if (valid[k]) rank[k] = m++;
    *n = O;
    *n_colours = stt.colours;
    for (uint32_t s = 0; s < O && order && s < cap; ++s) order[s] = rank[ord[s]];
    if (colour_offsets)
        for (uint32_t c = 0; c <= stt.colours && c < colour_cap; ++c) colour_offsets[c] = offs[c];
    return PXG_OK;
}
