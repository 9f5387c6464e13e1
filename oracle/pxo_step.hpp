// ORACLE -- test infrastructure only. Nothing in the product may include, link or run this.
//
// The world and its fixed step: physxcoopa/world.h:1278-1317 (step_fixed), the solver in
// physxcoopa/dynamics/solver.h:46-958 and the integrator in physxcoopa/dynamics/integrator.h.
#pragma once
#include "pxo_world.hpp"
#include "pxo_joints.hpp"
#include "pxo_pool.hpp"
#include "pxo_soft.hpp"

#include <algorithm>
#include <array>
#include <iterator>
#include <memory>
#include <set>

namespace pxo {

inline uint64_t pack_key(uint32_t ia, uint32_t ib, uint32_t fid) {  // solver.h:46-51
    uint32_t lo = std::min(ia, ib), hi = std::max(ia, ib);
    return (uint64_t(lo) << 40) ^ (uint64_t(hi) << 12) ^ uint64_t(fid);
}

template <class T> inline void orthonormal_basis(const V3<T>& n, V3<T>& t1, V3<T>& t2) {  // util/math.h:62-68
    T sign = n.z >= T(0) ? T(1) : T(-1);
    T a = T(-1) / (sign + n.z);
    T b = n.x * n.y * a;
    t1 = V3<T>(T(1) + sign * n.x * n.x * a, sign * b, -sign * n.x);
    t2 = V3<T>(b, sign + n.y * n.y * a, -n.y);
}

template <class T> inline void update_world_inertia(Body<T>& b) {  // integrator.h:57-68
    if (b.type != PXG_DYNAMIC) { b.iiw = M3<T>(T(0)); return; }
    M3<T> r = mat3_cast(b.q);
    M3<T> d(T(0));
    d.c[0].x = b.iil.x; d.c[1].y = b.iil.y; d.c[2].z = b.iil.z;
    b.iiw = mul(mul(r, d), transpose(r));
}

template <class T> inline void apply_constraints(Body<T>& b) {  // integrator.h:44-51
    if (b.constraints & PXG_FREEZE_POS_X) b.v.x = 0;
    if (b.constraints & PXG_FREEZE_POS_Y) b.v.y = 0;
    if (b.constraints & PXG_FREEZE_POS_Z) b.v.z = 0;
    if (b.constraints & PXG_FREEZE_ROT_X) b.w.x = 0;
    if (b.constraints & PXG_FREEZE_ROT_Y) b.w.y = 0;
    if (b.constraints & PXG_FREEZE_ROT_Z) b.w.z = 0;
}

template <class T> struct Config {
    uint32_t vel_iters = 16, relax_iters = 2, pos_iters = 4;
    T pos_corr = T(0.2f), slop = T(0.005f), max_corr = T(0.2f), rest_thresh = T(0.5f);
    T sleep_lin = T(0.01f), sleep_ang = T(0.02f), sleep_time = T(0.5f), margin = T(0.1f);
    uint32_t max_substeps = 8;
    T fixed_dt = T(1.0f / 60.0f), max_lin_vel = T(200.0f);
    bool exact_quat = false;
    V3<T> gravity{T(0), T(0), T(-9.81f)};
    bool residual_mode = false;
    T residual_tol = T(1e-4f);
    uint32_t residual_max = 256;
    uint32_t residual_every = 1;  // residual evaluated after every k-th pass and the last
    uint32_t cloth_substeps = 4, cloth_iterations = 1;  // util/config.h:139-144
};

template <class T> struct World {
    Config<T> cfg;
    std::vector<Body<T>> bodies;
    std::vector<Shape<T>> shapes;
    std::vector<Hull<T>> hulls;  // the hulls PXG_CONVEX_HULL shapes point at
    std::vector<Mesh> meshes;    // the cooked meshes PXG_TRIANGLE_MESH shapes point at
    std::vector<int32_t> proxy;      // per shape: tree node, -1 = none
    std::vector<char> in_static;     // per shape
    Tree<T> static_tree, dyn_tree;
    uint32_t layer_mask[32];
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    std::vector<Manifold<T>> manifolds;          // pair order (world.h:1647-1653)
    std::unordered_map<uint64_t, Cached> cache;  // solver.h:63-69
    // solve order for the next step: shape pairs (sa, sb) in the order the solver should visit
    // their manifolds; manifolds missing from it keep pair order after the listed ones.
    std::vector<std::pair<uint32_t, uint32_t>> order_hint;
    bool use_order_hint = false;
    // learned warm start: initial impulses for warm-start misses, keyed like the cache
    std::unordered_map<uint64_t, Cached> miss_table;
    bool use_miss_table = false;
    uint32_t brute_force = 0;
    // joints (world.h:2125) and the bodies whose contacts are switched off (world.h:2127, keys
    // lo << 32 | hi of the two body indices)
    std::vector<Joint<T>> joints;
    std::set<uint64_t> ignored;
    Atan2Mode atan2_mode = Atan2Mode::Libm;
    // cloths and hair (world.h:538-700): slot tables, a dead slot is not stepped or hashed
    std::vector<Cloth<T>> cloths;
    std::vector<char> cloth_alive;
    std::vector<StrandSet<T>> strand_sets;
    std::vector<char> strand_alive;
    // candidate colliders of a cloth or groom: true visits them in the reference's order (static
    // tree, then dynamic tree, each in its traversal order); false in shape slot order within each
    // tree (the order the GPU uses). The two differ only for a particle that meets two colliders of
    // one tree in one substep.
    bool collider_tree_order = true;
    T interpolation_alpha = T(0);
    // contact events (collision/contact_event.h:29-36): last step's confirmed shape pairs,
    // [0] solid and [1] trigger (world.h:2112-2113), and the queue callers drain (world.h:744)
    struct Event { uint32_t a, b, type, trigger; };
    std::vector<std::pair<uint32_t, uint32_t>> ev_prev[2];
    std::vector<Event> events;
    // diagnostics of the last step
    uint32_t vel_passes = 0;
    T last_residual = 0;
    std::vector<T> residual_trace;
    // per contact point after the last velocity pass (before the relax passes), when asked for:
    // {manifold in solve order, point, rn, rt1, rt2, lambda_n, lambda_t1, lambda_t2, v_n, v_t1,
    //  v_t2, target, separation, a active, b active, k_n}
    bool want_residual_detail = false;
    std::vector<std::array<double, 16>> residual_detail;
    uint64_t steps = 0;
    T accumulator = 0;
    // host threads for bounds, pair discovery and narrowphase (the reference's job-parallel
    // stages); none: serial, the reference's default (jobs_ == nullptr, world.h:2161)
    std::unique_ptr<Pool> pool;
    void set_threads(unsigned n) { pool.reset(n > 1 ? new Pool(n) : nullptr); }
    template <class Fn> void for_range(size_t count, Fn&& fn) {
        if (pool) pool->run(count, fn);
        else fn(size_t(0), count, 0u);
    }

    World() { for (int i = 0; i < 32; ++i) layer_mask[i] = 0xFFFFFFFFu; }

    bool should_collide(uint32_t a, uint32_t b) const { return (layer_mask[a & 31] & (1u << (b & 31))) != 0; }  // layer_matrix.h:28-30

    void set_margin(T m) { static_tree.margin = m; dyn_tree.margin = m; cfg.margin = m; }

    // world.h:1369-1382
    void create_proxies() {
        static_tree = Tree<T>(); dyn_tree = Tree<T>();
        static_tree.margin = cfg.margin; dyn_tree.margin = cfg.margin;
        proxy.assign(shapes.size(), -1);
        in_static.assign(shapes.size(), 0);
        for (size_t i = 0; i < shapes.size(); ++i) {
            const Shape<T>& s = shapes[i];
            if (!s.enabled) continue;
            const Body<T>& b = bodies[s.owner];
            AABB<T> bb = world_bounds(s, b.pos, b.q);
            if (b.type == PXG_STATIC) { proxy[i] = static_tree.create(bb, int32_t(i)); in_static[i] = 1; }
            else proxy[i] = dyn_tree.create(bb, int32_t(i));
        }
    }
    AABB<T> fat(size_t i) const { return in_static[i] ? static_tree.nodes[proxy[i]].box : dyn_tree.nodes[proxy[i]].box; }
    // Restores fat AABBs (e.g. from a GPU snapshot): rebuilds both trees with the given boxes.
    void set_fat(const std::vector<AABB<T>>& boxes) {
        static_tree = Tree<T>(); dyn_tree = Tree<T>();
        static_tree.margin = cfg.margin; dyn_tree.margin = cfg.margin;
        for (size_t i = 0; i < shapes.size(); ++i) {
            if (proxy[i] < 0) continue;
            Tree<T>& t = in_static[i] ? static_tree : dyn_tree;
            // create() expands by margin; undo by creating with the box shrunk then overwrite
            proxy[i] = t.create(boxes[i], int32_t(i));
            t.nodes[proxy[i]].box = boxes[i];
        }
        // refit internal nodes bottom-up so queries stay conservative over the overwritten leaves
        for (Tree<T>* t : {&static_tree, &dyn_tree}) {
            for (size_t i = 0; i < shapes.size(); ++i) {
                if (proxy[i] < 0 || (t == &static_tree) != bool(in_static[i])) continue;
                int32_t p = t->nodes[proxy[i]].parent;
                while (p >= 0) {
                    t->nodes[p].box = AABB<T>::merge(t->nodes[t->nodes[p].c1].box, t->nodes[t->nodes[p].c2].box);
                    p = t->nodes[p].parent;
                }
            }
        }
    }

    static T min_extent(const Shape<T>& s) {  // world.h:1411-1424
        if (s.type == PXG_CONVEX_HULL) return hull_min_extent(s);
        if (s.type == PXG_TRIANGLE_MESH) return std::numeric_limits<float>::max();
        if (s.type == PXG_SPHERE) return s.radius;
        if (s.type == PXG_BOX) return std::min(std::min(s.he.x, s.he.y), s.he.z);
        return std::min(s.cr, s.chh);
    }

    // sync_broadphase_ (world.h:1511-1518): tight bounds, and the dynamic tree's proxies moved
    std::vector<AABB<T>> sync_bounds(T h) {
        size_t ns = shapes.size();
        std::vector<AABB<T>> tight(ns);
        for_range(ns, [&](size_t begin, size_t end, unsigned) {  // world.h:1489-1499
            for (size_t i = begin; i < end; ++i) {
                if (proxy[i] < 0) continue;
                const Body<T>& b = bodies[shapes[i].owner];
                tight[i] = world_bounds(shapes[i], b.pos, b.q);
            }
        });
        for (size_t i = 0; i < ns; ++i) {
            if (proxy[i] < 0 || in_static[i]) continue;
            const Body<T>& b = bodies[shapes[i].owner];
            dyn_tree.move(proxy[i], tight[i], b.v * h);
        }
        return tight;
    }
    // world.h:1511-1518 + 1545-1577 + pair_cache.h:49-52
    void broadphase(T h) {
        size_t ns = shapes.size();
        std::vector<AABB<T>> tight = sync_bounds(h);
        pairs.clear();
        if (brute_force) {
            // the same rule, by brute force: (i, j) iff owners differ, not both static, and
            // tight(i) overlaps fat(j) or tight(j) overlaps fat(i)
            for (size_t i = 0; i < ns; ++i) {
                if (proxy[i] < 0) continue;
                for (size_t j = i + 1; j < ns; ++j) {
                    if (proxy[j] < 0) continue;
                    if (shapes[i].owner == shapes[j].owner) continue;
                    if (in_static[i] && in_static[j]) continue;
                    if (tight[i].overlaps(fat(j)) || tight[j].overlaps(fat(i))) pairs.push_back({uint32_t(i), uint32_t(j)});
                }
            }
            return;
        }
        // per-worker pair buckets (world.h:1545-1577), merged and sorted into the pair cache
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> buckets(pool ? pool->workers() : 1u);
        for_range(ns, [&](size_t begin, size_t end, unsigned worker) {
            auto& out = buckets[worker];
            for (size_t k = begin; k < end; ++k) {
                uint32_t i = uint32_t(k);
                if (proxy[i] < 0) continue;
                uint32_t oi = shapes[i].owner;
                dyn_tree.query(tight[i], [&](int32_t j) {
                    if (oi != shapes[j].owner) out.push_back({std::min(i, uint32_t(j)), std::max(i, uint32_t(j))});
                });
                if (!in_static[i]) {
                    static_tree.query(tight[i], [&](int32_t j) {
                        if (oi != shapes[j].owner) out.push_back({std::min(i, uint32_t(j)), std::max(i, uint32_t(j))});
                    });
                }
            }
        });
        for (auto& b : buckets) pairs.insert(pairs.end(), b.begin(), b.end());
        std::sort(pairs.begin(), pairs.end());
        pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
        // PairCache::add drops x == y (pair_cache.h:43-46); owners differ so shapes differ too
    }

    static uint64_t body_pair_key(uint32_t a, uint32_t b) {  // world.h:1657-1660
        return (uint64_t(std::min(a, b)) << 32) | std::max(a, b);
    }
    bool joint_blocks(uint32_t i, uint32_t j) const {  // world.h:1686-1694
        for (const auto& jt : joints) {
            if (!jt.valid || jt.collide_connected) continue;
            if ((jt.a == i && jt.b == j) || (jt.a == j && jt.b == i)) return true;
        }
        return false;
    }
    JointSolver<T, Body<T>> joint_solver() { return JointSolver<T, Body<T>>{bodies, joints, atan2_mode}; }

    // world.h:1598-1654
    // per pair in parallel (world.h:1598-1645), then compacted in pair order (1647-1653)
    void narrowphase(T h) {
        manifolds.clear();
        std::vector<Manifold<T>> slot(pairs.size());
        for_range(pairs.size(), [&](size_t begin, size_t end, unsigned) {
            for (size_t k = begin; k < end; ++k) {
                uint32_t i = pairs[k].first, j = pairs[k].second;
                const Shape<T>& si = shapes[i];
                const Shape<T>& sj = shapes[j];
                uint32_t ii = si.owner, ij = sj.owner;
                const Body<T>& bi = bodies[ii];
                const Body<T>& bj = bodies[ij];
                if (!bi.alive || !bj.alive) continue;
                bool either_dyn = bi.type == PXG_DYNAMIC || bj.type == PXG_DYNAMIC;
                if (!either_dyn) continue;
                if (!should_collide(si.layer, sj.layer)) continue;
                if (joint_blocks(ii, ij)) continue;
                if (!ignored.empty() && ignored.count(body_pair_key(ii, ij))) continue;
                T ei = min_extent(si), ej = min_extent(sj);
                bool spec = (length(bi.v) * h > ei) || (length(bj.v) * h > ej);
                Manifold<T>& m = slot[k];
                generate_contacts(ii, si, bi.pos, bi.q, ij, sj, bj.pos, bj.q, m, spec);
                if (!m.valid) continue;
                m.sa = i; m.sb = j;
            }
        });
        for (auto& m : slot)
            if (m.valid) manifolds.push_back(m);
    }

    // ---------------- solver (dynamics/solver.h) ----------------
    static bool active(const Body<T>& b) { return b.type == PXG_DYNAMIC && b.awake; }

    static void apply_pair(Body<T>& a, Body<T>& b, const V3<T>& p, const V3<T>& J) {  // solver.h:86-95
        if (active(a)) { a.v -= J * a.inv_mass; a.w -= mul(a.iiw, cross(p - a.pos, J)); }
        if (active(b)) { b.v += J * b.inv_mass; b.w += mul(b.iiw, cross(p - b.pos, J)); }
    }

    void capture_restitution(std::vector<Manifold<T>*>& ms) {  // solver.h:121-137
        for (Manifold<T>* mp : ms) {
            Manifold<T>& m = *mp;
            if (!m.valid || m.trigger) continue;
            Body<T>& a = bodies[m.a];
            Body<T>& b = bodies[m.b];
            for (int i = 0; i < m.count; ++i) {
                Contact<T>& p = m.p[i];
                V3<T> ra = p.pos - a.pos, rb = p.pos - b.pos;
                V3<T> va = a.v + cross(a.w, ra), vb = b.v + cross(b.w, rb);
                T vap = dot(vb - va, m.n);
                p.rbias = (-vap > cfg.rest_thresh) ? -m.restitution * vap : T(0);
            }
        }
    }

    void warm_start(std::vector<Manifold<T>*>& ms, uint32_t& predicted) {  // solver.h:144-175
        for (Manifold<T>* mp : ms) {
            Manifold<T>& m = *mp;
            if (!m.valid || m.trigger) continue;
            Body<T>& a = bodies[m.a];
            Body<T>& b = bodies[m.b];
            if (!active(a) && !active(b)) continue;
            V3<T> t1, t2;
            orthonormal_basis(m.n, t1, t2);
            for (int i = 0; i < m.count; ++i) {
                Contact<T>& p = m.p[i];
                uint64_t key = pack_key(m.a, m.b, p.fid);
                auto it = cache.find(key);
                if (it != cache.end()) {
                    p.ni = T(it->second.n); p.ti[0] = T(it->second.t1); p.ti[1] = T(it->second.t2);
                } else if (use_miss_table && miss_table.count(key)) {
                    const Cached& c = miss_table[key];
                    p.ni = T(c.n); p.ti[0] = T(c.t1); p.ti[1] = T(c.t2);
                    ++predicted;
                } else {
                    p.ni = 0; p.ti[0] = 0; p.ti[1] = 0;
                }
                V3<T> J = m.n * p.ni + t1 * p.ti[0] + t2 * p.ti[1];
                if (dot(J, J) > T(0)) apply_pair(a, b, p.pos, J);
            }
        }
    }

    // solver.h:212-281; returns nothing, residual computed separately
    void velocity_pass(std::vector<Manifold<T>*>& ms, bool use_rest, uint32_t offset, T h) {
        for (Manifold<T>* mp : ms) {
            Manifold<T>& m = *mp;
            if (!m.valid || m.trigger) continue;
            Body<T>& a = bodies[m.a];
            Body<T>& b = bodies[m.b];
            bool ad = active(a), bd = active(b);
            if (!ad && !bd) continue;
            T ima = ad ? a.inv_mass : T(0), imb = bd ? b.inv_mass : T(0);
            M3<T> Ia = ad ? a.iiw : M3<T>(T(0));
            M3<T> Ib = bd ? b.iiw : M3<T>(T(0));
            V3<T> t1, t2;
            orthonormal_basis(m.n, t1, t2);
            for (int k = 0; k < m.count; ++k) {
                int i = int((uint32_t(k) + offset) % uint32_t(m.count));
                Contact<T>& p = m.p[i];
                V3<T> ra = p.pos - a.pos, rb = p.pos - b.pos;
                {
                    V3<T> ran = cross(ra, m.n), rbn = cross(rb, m.n);
                    T kn = ima + imb + dot(ran, mul(Ia, ran)) + dot(rbn, mul(Ib, rbn));
                    if (kn > k_epsilon<T>()) {
                        V3<T> va = a.v + cross(a.w, ra), vb = b.v + cross(b.w, rb);
                        T vn = dot(vb - va, m.n);
                        T target = p.sep > T(0) ? -p.sep / h : (use_rest ? p.rbias : T(0));
                        T lambda = -(vn - target) / kn;
                        T ni = gmax(p.ni + lambda, T(0));
                        T delta = ni - p.ni;
                        p.ni = ni;
                        apply_pair(a, b, p.pos, m.n * delta);
                    }
                }
                const V3<T>* tg[2] = {&t1, &t2};
                for (int t = 0; t < 2; ++t) {
                    const V3<T>& dir = *tg[t];
                    V3<T> rat = cross(ra, dir), rbt = cross(rb, dir);
                    T kt = ima + imb + dot(rat, mul(Ia, rat)) + dot(rbt, mul(Ib, rbt));
                    if (kt <= k_epsilon<T>()) continue;
                    V3<T> va = a.v + cross(a.w, ra), vb = b.v + cross(b.w, rb);
                    T vt = dot(vb - va, dir);
                    T lt = -vt / kt;
                    T maxf = m.friction * p.ni;
                    T old = p.ti[t];
                    T nt = sclamp(old + lt, -maxf, maxf);
                    T delta = nt - old;
                    p.ti[t] = nt;
                    apply_pair(a, b, p.pos, dir * delta);
                }
            }
        }
    }

    // Velocity residual over active contact rows (m/s): for the normal row the complementarity
    // violation of (lambda_n, v_n - target), for each friction row the violation of the Coulomb
    // cone's stick/slip conditions. Evaluated on the current velocities and impulses.
    T residual(std::vector<Manifold<T>*>& ms, bool use_rest, T h, std::vector<std::array<double, 16>>* detail = nullptr) {
        T r = 0;
        if (detail) detail->clear();
        for (size_t mi = 0; mi < ms.size(); ++mi) {
            Manifold<T>& m = *ms[mi];
            if (!m.valid || m.trigger) continue;
            Body<T>& a = bodies[m.a];
            Body<T>& b = bodies[m.b];
            if (!active(a) && !active(b)) continue;
            V3<T> t1, t2;
            orthonormal_basis(m.n, t1, t2);
            for (int i = 0; i < m.count; ++i) {
                Contact<T>& p = m.p[i];
                V3<T> ra = p.pos - a.pos, rb = p.pos - b.pos;
                V3<T> va = a.v + cross(a.w, ra), vb = b.v + cross(b.w, rb);
                V3<T> dv = vb - va;
                T vn = dot(dv, m.n);
                T target = p.sep > T(0) ? -p.sep / h : (use_rest ? p.rbias : T(0));
                T rn = p.ni > T(0) ? std::abs(vn - target) : gmax(T(0), target - vn);
                r = gmax(r, rn);
                T maxf = m.friction * p.ni;
                const V3<T>* tg[2] = {&t1, &t2};
                T rts[2], vts[2];
                for (int t = 0; t < 2; ++t) {
                    T vt = dot(dv, *tg[t]);
                    // stick: |lambda_t| < mu*lambda_n needs v_t = 0; slip: the friction impulse
                    // (applied +lambda_t to B) must oppose v_t, i.e. v_t * sign(lambda_t) <= 0
                    T rt = T(0);
                    if (maxf > T(0)) rt = std::abs(p.ti[t]) < maxf ? std::abs(vt) : gmax(T(0), (p.ti[t] < T(0) ? -vt : vt));
                    r = gmax(r, rt);
                    rts[t] = rt; vts[t] = vt;
                }
                if (detail) {
                    bool ad = active(a), bd = active(b);
                    T ima = ad ? a.inv_mass : T(0), imb = bd ? b.inv_mass : T(0);
                    M3<T> Ia = ad ? a.iiw : M3<T>(T(0));
                    M3<T> Ib = bd ? b.iiw : M3<T>(T(0));
                    V3<T> ran = cross(ra, m.n), rbn = cross(rb, m.n);
                    T kn = ima + imb + dot(ran, mul(Ia, ran)) + dot(rbn, mul(Ib, rbn));
                    detail->push_back({double(mi), double(i), double(rn), double(rts[0]), double(rts[1]), double(p.ni),
                                       double(p.ti[0]), double(p.ti[1]), double(vn), double(vts[0]), double(vts[1]),
                                       double(target), double(p.sep), double(ad), double(bd), double(kn)});
                }
            }
        }
        return r;
    }

    void solve_position(std::vector<Manifold<T>*>& ms) {  // solver.h:293-328
        std::vector<V3<T>> ref(bodies.size());
        for (size_t i = 0; i < bodies.size(); ++i) ref[i] = bodies[i].pos;
        for (uint32_t it = 0; it < cfg.pos_iters; ++it) {
            for (Manifold<T>* mp : ms) {
                Manifold<T>& m = *mp;
                if (!m.valid || m.trigger) continue;
                Body<T>& a = bodies[m.a];
                Body<T>& b = bodies[m.b];
                bool ad = active(a), bd = active(b);
                if (!ad && !bd) continue;
                T ima = ad ? a.inv_mass : T(0), imb = bd ? b.inv_mass : T(0);
                T k = ima + imb;
                if (k <= k_epsilon<T>()) continue;
                V3<T> da = a.pos - ref[m.a], db = b.pos - ref[m.b];
                for (int i = 0; i < m.count; ++i) {
                    Contact<T>& p = m.p[i];
                    T cur = p.pen - dot(db - da, m.n);
                    T mag = gmax(T(0), cur - cfg.slop) * cfg.pos_corr;
                    mag = gmin(mag, cfg.max_corr);
                    if (mag <= T(0)) continue;
                    V3<T> corr = m.n * (mag / k);
                    if (ad) { a.pos -= corr * ima; da = a.pos - ref[m.a]; }
                    if (bd) { b.pos += corr * imb; db = b.pos - ref[m.b]; }
                }
            }
        }
    }

    // solver.h:767-781. The reference inserts in its manifold order (pair order); with a solve
    // order hint the insertion still follows pair order, so a key collision resolves the same way.
    void rebuild_cache() {
        cache.clear();
        for (Manifold<T>& m : manifolds) {
            if (!m.valid || m.trigger) continue;
            for (int i = 0; i < m.count; ++i) {
                Cached c;
                c.n = double(m.p[i].ni); c.t1 = double(m.p[i].ti[0]); c.t2 = double(m.p[i].ti[1]);
                cache[pack_key(m.a, m.b, m.p[i].fid)] = c;
            }
        }
    }

    void wake_if_kin_moving(Body<T>& kin, Body<T>& dyn) {  // solver.h:787-795
        if (kin.type != PXG_KINEMATIC || dyn.type != PXG_DYNAMIC || dyn.awake) return;
        bool lm = dot(kin.v, kin.v) > cfg.sleep_lin * cfg.sleep_lin;
        bool am = dot(kin.w, kin.w) > cfg.sleep_ang * cfg.sleep_ang;
        if (lm || am) { dyn.awake = true; dyn.sleep_timer = 0; }
    }

    void islands_and_sleep(T h) {  // solver.h:808-876 + island.h:25-49
        std::vector<uint32_t> parent(bodies.size());
        for (size_t i = 0; i < parent.size(); ++i) parent[i] = uint32_t(i);
        auto find = [&](uint32_t i) {
            while (parent[i] != i) { parent[i] = parent[parent[i]]; i = parent[i]; }
            return i;
        };
        for (auto& m : manifolds) {
            if (!m.valid) continue;
            if (bodies[m.a].type == PXG_DYNAMIC && bodies[m.b].type == PXG_DYNAMIC) {
                uint32_t ra = find(m.a), rb = find(m.b);
                if (ra != rb) parent[ra] = rb;
            }
        }
        for (auto& j : joints) {
            if (!j.valid || !j.enabled) continue;
            if (bodies[j.a].type == PXG_DYNAMIC && bodies[j.b].type == PXG_DYNAMIC) {
                uint32_t ra = find(j.a), rb = find(j.b);
                if (ra != rb) parent[ra] = rb;
            }
        }
        for (auto& b : bodies) {
            if (!b.alive || b.type != PXG_DYNAMIC || !b.awake) continue;
            T l2 = dot(b.v, b.v), a2 = dot(b.w, b.w);
            if (l2 < cfg.sleep_lin * cfg.sleep_lin && a2 < cfg.sleep_ang * cfg.sleep_ang) b.sleep_timer += h;
            else b.sleep_timer = 0;
        }
        for (auto& m : manifolds) {
            if (!m.valid) continue;
            wake_if_kin_moving(bodies[m.a], bodies[m.b]);
            wake_if_kin_moving(bodies[m.b], bodies[m.a]);
        }
        for (auto& j : joints) {
            if (!j.valid || !j.enabled) continue;
            wake_if_kin_moving(bodies[j.a], bodies[j.b]);
            wake_if_kin_moving(bodies[j.b], bodies[j.a]);
        }
        std::unordered_map<uint32_t, T> mint;
        for (uint32_t i = 0; i < bodies.size(); ++i) {
            Body<T>& b = bodies[i];
            if (!b.alive || b.type != PXG_DYNAMIC || !b.awake) continue;
            uint32_t r = find(i);
            auto it = mint.find(r);
            if (it == mint.end() || b.sleep_timer < it->second) mint[r] = b.sleep_timer;
        }
        for (uint32_t i = 0; i < bodies.size(); ++i) {
            Body<T>& b = bodies[i];
            if (!b.alive || b.type != PXG_DYNAMIC || !b.awake) continue;
            if (mint[find(i)] >= cfg.sleep_time) { b.awake = false; b.v = V3<T>(0); b.w = V3<T>(0); }
        }
    }

    // The solve order: manifolds listed in order_hint first (in that order), then the rest in
    // pair order. With no hint this is the reference order (manifolds_ as compacted).
    std::vector<Manifold<T>*> solve_order() {
        std::vector<Manifold<T>*> ms;
        ms.reserve(manifolds.size());
        if (!use_order_hint) {
            for (auto& m : manifolds) ms.push_back(&m);
            return ms;
        }
        std::unordered_map<uint64_t, size_t> idx;
        for (size_t i = 0; i < manifolds.size(); ++i) idx[(uint64_t(manifolds[i].sa) << 32) | manifolds[i].sb] = i;
        std::vector<char> used(manifolds.size(), 0);
        for (auto& pr : order_hint) {
            auto it = idx.find((uint64_t(pr.first) << 32) | pr.second);
            if (it == idx.end() || used[it->second]) continue;
            used[it->second] = 1;
            ms.push_back(&manifolds[it->second]);
        }
        for (size_t i = 0; i < manifolds.size(); ++i) if (!used[i]) ms.push_back(&manifolds[i]);
        return ms;
    }

    uint32_t last_predicted = 0;

    // dynamics/solver.h:914-958
    void solve(T h) {
        JointSolver<T, Body<T>> js = joint_solver();
        js.wake_jointed();
        std::vector<Manifold<T>*> ms = solve_order();
        capture_restitution(ms);
        last_predicted = 0;
        warm_start(ms, last_predicted);
        js.warm_start();
        residual_trace.clear();
        if (cfg.residual_mode) {
            // at least one pass; the residual after every k-th pass and after the cap's last
            uint32_t i = 0;
            const uint32_t every = std::max(1u, cfg.residual_every);
            T r = residual(ms, false, h);
            while (i < cfg.residual_max) {
                velocity_pass(ms, false, i, h);
                js.velocity_pass(h, (i & 1u) != 0);
                ++i;
                if (i % every != 0u && i != cfg.residual_max) continue;
                r = residual(ms, false, h);
                residual_trace.push_back(r);
                if (r <= cfg.residual_tol) break;
            }
            vel_passes = i;
            last_residual = r;
            if (want_residual_detail) residual(ms, false, h, &residual_detail);
        } else {
            for (uint32_t i = 0; i < cfg.vel_iters; ++i) {
                velocity_pass(ms, false, i, h);
                js.velocity_pass(h, (i & 1u) != 0);
            }
            vel_passes = cfg.vel_iters;
            if (want_residual_detail) residual(ms, false, h, &residual_detail);
        }
        for (uint32_t i = 0; i < cfg.relax_iters; ++i) {
            velocity_pass(ms, true, i, h);
            js.velocity_pass(h, (i & 1u) != 0);
        }
        if (!cfg.residual_mode) last_residual = residual(ms, cfg.relax_iters > 0, h);
        solve_position(ms);
        js.position_pass(cfg);
        rebuild_cache();
        islands_and_sleep(h);
    }

    void integrate_forces_all(T h) {  // world.h:1283-1295, integrator.h:31-35 + 78-94
        for (auto& b : bodies) {
            if (!b.alive) continue;
            if (b.type == PXG_DYNAMIC && b.awake && b.use_gravity) b.force += cfg.gravity * b.mass;
        }
        for (auto& b : bodies) {
            if (!b.alive) continue;
            if (b.type != PXG_DYNAMIC || !b.awake) { b.force = V3<T>(0); b.torque = V3<T>(0); continue; }
            b.v += b.force * b.inv_mass * h;
            b.w += mul(b.iiw, b.torque) * h;
            b.v *= T(1) / (T(1) + h * b.ldrag);
            b.w *= T(1) / (T(1) + h * b.adrag);
            apply_constraints(b);
            b.force = V3<T>(0); b.torque = V3<T>(0);
        }
    }

    void integrate_velocities_all(T h) {  // world.h:1303-1309, integrator.h:109-133
        for (auto& b : bodies) {
            if (!b.alive) continue;
            b.prev_pos = b.pos;
            b.prev_q = b.q;
            if (b.type == PXG_DYNAMIC && b.awake) {
                T s2 = dot(b.v, b.v);
                T m2 = cfg.max_lin_vel * cfg.max_lin_vel;
                if (s2 > m2 && s2 > k_epsilon<T>()) b.v *= cfg.max_lin_vel / std::sqrt(s2);
                b.pos += b.v * h;
                if (cfg.exact_quat) {
                    V3<T> w = b.w;
                    T angle = length(w) * h;
                    if (angle > k_epsilon<T>()) {
                        V3<T> ax = w / length(w);
                        T half = angle * T(0.5f);
                        T s = std::sin(half);
                        Q<T> dq = Q<T>::wxyz(std::cos(half), ax.x * s, ax.y * s, ax.z * s);
                        b.q = qnormalize(qmul(dq, b.q));
                    }
                } else {
                    Q<T> wq = Q<T>::wxyz(0, b.w.x, b.w.y, b.w.z);
                    Q<T> dq = qmul(wq, b.q);
                    b.q = qnormalize(qadd(b.q, qscale(dq, T(0.5f) * h)));
                }
            }
            update_world_inertia(b);
        }
    }

    // world.h:1964-1978 (the caches roll over and refill) and 1998-2018 (emit_contact_events_):
    // per cache, PairCache::classify (pair_cache.h:63-82) as sorted-set differences, pushed as
    // all enters, then all stays, then all exits; the solid cache before the trigger cache.
    void emit_events() {
        typedef std::pair<uint32_t, uint32_t> PP;
        std::vector<PP> cur[2];
        for (const auto& m : manifolds) cur[m.trigger ? 1 : 0].push_back(PP(m.sa, m.sb));
        for (int k = 0; k < 2; ++k) {
            std::sort(cur[k].begin(), cur[k].end());  // PairCache::finalize
            cur[k].erase(std::unique(cur[k].begin(), cur[k].end()), cur[k].end());
            std::vector<PP> added, stayed, removed;
            std::set_difference(cur[k].begin(), cur[k].end(), ev_prev[k].begin(), ev_prev[k].end(), std::back_inserter(added));
            std::set_difference(ev_prev[k].begin(), ev_prev[k].end(), cur[k].begin(), cur[k].end(), std::back_inserter(removed));
            std::set_difference(cur[k].begin(), cur[k].end(), added.begin(), added.end(), std::back_inserter(stayed));
            auto push = [&](const PP& p, uint32_t type) {
                if (p.first >= shapes.size() || p.second >= shapes.size()) return;
                if (type != 2u && (!shapes[p.first].enabled || !shapes[p.second].enabled)) return;
                uint32_t oa = shapes[p.first].owner, ob = shapes[p.second].owner;
                if (oa >= bodies.size() || ob >= bodies.size() || !bodies[oa].alive || !bodies[ob].alive) return;
                events.push_back(Event{oa, ob, type, uint32_t(k)});
            };
            for (const PP& p : added) push(p, 0u);
            for (const PP& p : stayed) push(p, 1u);
            for (const PP& p : removed) push(p, 2u);
            ev_prev[k].swap(cur[k]);
        }
    }

    void step_phase_collide(T h) {
        integrate_forces_all(h);
        broadphase(h);
        narrowphase(h);
        emit_events();
    }
    void step_phase_solve(T h) {
        solve(h);
        integrate_velocities_all(h);
        if (!joints.empty()) {  // world.h:1310-1314, solver.h:755-761
            JointSolver<T, Body<T>> js = joint_solver();
            for (uint32_t it = 0; it < std::max(1u, cfg.pos_iters); ++it) js.position_pass(cfg);
            for (auto& b : bodies) if (b.alive) update_world_inertia(b);
        }
        ++steps;
    }
    void step_fixed(T h) { step_phase_collide(h); step_phase_solve(h); }  // world.h:1278-1317

    // step() (world.h:1229-1257): fixed substeps from the accumulator, then cloth and hair once
    // per frame (the broadphase synced first when no substep ran). Returns the substeps run.
    uint32_t step(T dt) {
        events.clear();
        const T frame = std::min(dt, T(0.25f));
        accumulator += frame;
        const T h = cfg.fixed_dt;
        uint32_t n = 0;
        while (accumulator >= h && n < cfg.max_substeps) {
            step_fixed(h);
            accumulator -= h;
            ++n;
        }
        interpolation_alpha = accumulator / h;
        const bool soft = !cloths.empty() || !strand_sets.empty();
        if (soft && frame > T(0)) {
            if (n == 0) sync_bounds(frame);
            if (!cloths.empty()) step_cloths(frame);
            if (!strand_sets.empty()) step_strands(frame);
        }
        return n;
    }
    bool body_valid(bool has, uint32_t index, uint32_t generation) const {
        return has && index < bodies.size() && bodies[index].alive && bodies[index].generation == generation;
    }
    // the colliders a cloth or groom meets (world.h:1778-1810, 1880-1912): enabled, non-trigger
    // shapes of live bodies that the layer matrix lets collide with `layer` and whose world
    // bounds overlap `query`; `moving` is set when one belongs to a moving non-static body
    void add_collider(uint32_t slot, uint32_t layer, const AABB<T>& query, std::vector<Collider<T>>& out, bool& moving) {
        if (slot >= shapes.size()) return;
        const Shape<T>& sh = shapes[slot];
        if (!sh.enabled || sh.trigger) return;
        if (!should_collide(layer, sh.layer)) return;
        const Body<T>& b = bodies[sh.owner];
        if (!b.alive) return;
        Collider<T> c;
        c.shape = &sh;
        c.pos = b.pos;
        c.rot = b.q;
        c.bounds = world_bounds(sh, b.pos, b.q);
        if (!c.bounds.overlaps(query)) return;
        out.push_back(c);
        if (b.type != PXG_STATIC && (dot(b.v, b.v) > k_epsilon<T>() || dot(b.w, b.w) > k_epsilon<T>())) moving = true;
    }
    void gather_colliders(uint32_t layer, const AABB<T>& query, std::vector<Collider<T>>& out, bool& moving) {
        out.clear();
        if (collider_tree_order) {
            static_tree.query(query, [&](int32_t j) { add_collider(uint32_t(j), layer, query, out, moving); });
            dyn_tree.query(query, [&](int32_t j) { add_collider(uint32_t(j), layer, query, out, moving); });
            return;
        }
        for (int pass = 0; pass < 2; ++pass)
            for (size_t j = 0; j < shapes.size(); ++j) {
                if (proxy[j] < 0 || bool(in_static[j]) != (pass == 0)) continue;
                if (!fat(j).overlaps(query)) continue;
                add_collider(uint32_t(j), layer, query, out, moving);
            }
    }
    uint32_t soft_substeps(T dt) const {  // world.h:1753-1757
        const T internal = std::max(cfg.fixed_dt / T(std::max(1u, cfg.cloth_substeps)), k_epsilon<T>());
        const long r = std::lround(dt / internal);
        return uint32_t(std::clamp<long>(r, 1, 16));
    }
    // step_cloths_ (world.h:1749-1847)
    void step_cloths(T dt) {
        const uint32_t fs = soft_substeps(dt);
        std::vector<Collider<T>> cols;
        for (size_t ci = 0; ci < cloths.size(); ++ci) {
            if (!cloth_alive[ci]) continue;
            Cloth<T>& c = cloths[ci];
            if (!c.enabled || c.particles.empty()) continue;
            bool anchor_moved = false;
            for (CAnchor<T>& a : c.anchors) {
                const bool live = body_valid(a.has_body, a.body, a.generation);
                const V3<T> target = live ? bodies[a.body].pos + qrot(bodies[a.body].q, a.local) : a.world;
                const V3<T> d = target - a.world;
                if (dot(d, d) > k_epsilon<T>()) anchor_moved = true;
                a.world = target;
            }
            const T travel = max_particle_speed(c.particles) * dt;
            const AABB<T> query = c.bounds.expand(c.params.thickness + travel + cfg.margin);
            bool moving = false;
            gather_colliders(c.params.layer, query, cols, moving);
            if (anchor_moved || moving) {
                c.awake = true;
                c.sleep_timer = T(0);
            }
            if (!c.awake) {
                for (CAnchor<T>& a : c.anchors) a.prev_world = a.world;
                continue;
            }
            solve_cloth(c, cols, cfg.gravity, fs, cfg.cloth_iterations, dt, atan2_mode);
            if (max_particle_speed(c.particles) < c.params.sleep_threshold) {
                c.sleep_timer += dt;
                if (c.sleep_timer >= c.params.sleep_time) {
                    c.awake = false;
                    for (CParticle<T>& p : c.particles) p.vel = V3<T>(T(0));
                }
            } else {
                c.sleep_timer = T(0);
            }
        }
    }
    // step_strands_ (world.h:1861-1957)
    void step_strands(T dt) {
        const uint32_t fs = soft_substeps(dt);
        std::vector<Collider<T>> cols;
        for (size_t si = 0; si < strand_sets.size(); ++si) {
            if (!strand_alive[si]) continue;
            StrandSet<T>& set = strand_sets[si];
            if (!set.enabled || set.empty()) continue;
            bool root_moved = false;
            for (const Strand<T>& st : set.strands) {
                const V3<T> d = st.root - st.prev_root;
                if (dot(d, d) > k_epsilon<T>()) { root_moved = true; break; }
            }
            const T travel = max_particle_speed(set.particles) * dt;
            const AABB<T> query = set.bounds.expand(set.params.thickness + travel + cfg.margin);
            bool moving = false;
            if (set.collide_bodies.empty()) {
                gather_colliders(set.params.layer, query, cols, moving);
            } else {
                cols.clear();
                for (const auto& id : set.collide_bodies) {
                    if (!body_valid(true, id.first, id.second)) continue;
                    for (size_t j = 0; j < shapes.size(); ++j)
                        if (shapes[j].owner == id.first) add_collider(uint32_t(j), set.params.layer, query, cols, moving);
                }
            }
            if (root_moved || moving) {
                set.awake = true;
                set.sleep_timer = T(0);
            }
            if (!set.awake) {
                for (Strand<T>& st : set.strands) {
                    st.prev_root = st.root;
                    st.prev_root_q = st.root_q;
                }
                continue;
            }
            solve_strands(set, cols, cfg.gravity, fs, cfg.cloth_iterations, dt, atan2_mode);
            if (max_particle_speed(set.particles) < set.params.sleep_threshold) {
                set.sleep_timer += dt;
                if (set.sleep_timer >= set.params.sleep_time) {
                    set.awake = false;
                    for (CParticle<T>& p : set.particles) p.vel = V3<T>(T(0));
                }
            } else {
                set.sleep_timer = T(0);
            }
        }
    }

    uint64_t state_hash() const {  // world.h:1334-1359 (FP32 layout)
        uint64_t hsh = 1469598103934665603ull;
        auto mix = [&](const void* p, size_t n) {
            const uint8_t* c = static_cast<const uint8_t*>(p);
            for (size_t i = 0; i < n; ++i) { hsh ^= c[i]; hsh *= 1099511628211ull; }
        };
        for (const auto& b : bodies) {
            if (!b.alive) continue;
            float p[3] = {float(b.pos.x), float(b.pos.y), float(b.pos.z)};
            float q[4] = {float(b.q.x), float(b.q.y), float(b.q.z), float(b.q.w)};
            float v[3] = {float(b.v.x), float(b.v.y), float(b.v.z)};
            float w[3] = {float(b.w.x), float(b.w.y), float(b.w.z)};
            mix(p, 12); mix(q, 16); mix(v, 12); mix(w, 12);
        }
        for (size_t ci = 0; ci < cloths.size(); ++ci) {  // world.h:1352-1357
            if (!cloth_alive[ci]) continue;
            for (const CParticle<T>& pt : cloths[ci].particles) {
                float p[3] = {float(pt.pos.x), float(pt.pos.y), float(pt.pos.z)};
                float v[3] = {float(pt.vel.x), float(pt.vel.y), float(pt.vel.z)};
                mix(p, 12); mix(v, 12);
            }
        }
        return hsh;
    }
};

}  // namespace pxo
