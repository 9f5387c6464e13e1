// C++ facade over the pxg C ABI with the reference's PhysicsWorld interface
// (physxcoopa/world.h). A caller of physxcoopa::PhysicsWorld switches by including this header
// and linking libpxg.so; bodies and shapes are the ABI records of pxg.h.
//
// State model: the GPU owns body state between steps. The host keeps a mirror that is pulled
// on demand (get_body, body_count, ...) and pushed before the next step when it was handed out
// for writing or the body set changed. Errors follow the reference's conventions: stale handles
// are no-ops or nullptr (world.h:340-364); a failure inside the library throws
// std::runtime_error("[physxcoopa] ...") (util/error.h:24-26).
#pragma once

#include "pxg.h"

#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace pxg {

// Generational handle (dynamics/body.h:61-77).
struct BodyId {
    static constexpr uint32_t kInvalid = 0xFFFFFFFFu;
    uint32_t index = kInvalid;
    uint32_t generation = 0;
    bool is_valid() const { return index != kInvalid; }
    bool operator==(const BodyId& o) const { return index == o.index && generation == o.generation; }
    bool operator!=(const BodyId& o) const { return !(*this == o); }
};

// 32 x 32 symmetric collision matrix, all pairs enabled by default (broadphase/layer_matrix.h:21-47).
class LayerMatrix {
public:
    LayerMatrix() {
        for (auto& m : mask_) m = 0xFFFFFFFFu;
    }
    bool should_collide(uint32_t a, uint32_t b) const { return (mask_[a & 31u] >> (b & 31u)) & 1u; }
    void set_layer_collision(uint32_t a, uint32_t b, bool collide) {
        a &= 31u;
        b &= 31u;
        if (collide) {
            mask_[a] |= 1u << b;
            mask_[b] |= 1u << a;
        } else {
            mask_[a] &= ~(1u << b);
            mask_[b] &= ~(1u << a);
        }
        ++version_;
    }
    const uint32_t* rows() const { return mask_; }
    uint64_t version() const { return version_; }

private:
    uint32_t mask_[32];
    uint64_t version_ = 0;
};

// The slot list the reference's on_substep uses (coopa::event::Signal<PhysicsWorld&, float>).
template <class... Args>
class Signal {
public:
    using Slot = std::function<void(Args...)>;
    uint64_t connect(Slot s) {
        slots_.emplace_back(++next_, std::move(s));
        return next_;
    }
    void disconnect(uint64_t id) {
        for (size_t i = 0; i < slots_.size(); ++i)
            if (slots_[i].first == id) {
                slots_.erase(slots_.begin() + static_cast<std::ptrdiff_t>(i));
                return;
            }
    }
    bool empty() const { return slots_.empty(); }
    void emit(Args... a) {
        for (auto& s : slots_) s.second(a...);
    }

private:
    std::vector<std::pair<uint64_t, Slot>> slots_;
    uint64_t next_ = 0;
};

[[noreturn]] inline void raise_error(const std::string& what) {
    throw std::runtime_error("[physxcoopa] " + what + ": " + pxg_last_error());
}
inline void check(int status, const char* what) {
    if (status != PXG_OK) raise_error(what);
}

inline pxg_config default_config() {
    pxg_config c;
    pxg_config_default(&c);
    return c;
}
// A disabled shape: the body takes part in dynamics but never in collision (world.h:85-95).
inline pxg_shape no_shape() {
    pxg_shape s;
    std::memset(&s, 0, sizeof(s));
    s.local_rotation[3] = 1.0f;
    return s;
}

class PhysicsWorld {
public:
    // Fired once per step_fixed() after force integration and contact generation, before the
    // solve (world.h:70-78, 1299). Slots may change velocities and forces through get_body().
    Signal<PhysicsWorld&, float> on_substep;

    explicit PhysicsWorld(const pxg_config& config = default_config()) : cfg_(config) {
        // device buffers are sized once per handle: reserve room to add bodies later
        if (cfg_.max_bodies == 0) cfg_.max_bodies = 256;
        if (cfg_.max_shapes == 0) cfg_.max_shapes = cfg_.max_bodies;
        check(pxg_create(&cfg_, &w_), "pxg_create");
        applied_cfg_ = cfg_;
    }
    ~PhysicsWorld() {
        if (w_) pxg_destroy(w_);
    }
    PhysicsWorld(const PhysicsWorld&) = delete;
    PhysicsWorld& operator=(const PhysicsWorld&) = delete;

    // --- body lifecycle (world.h:95-143, 340-364) ---
    // From an on_substep slot the insertion is deferred to the top of the next step_fixed() and
    // the returned handle is invalid (world.h:281-300).
    BodyId add_body(const pxg_body& initial, const pxg_shape& shape = no_shape()) {
        if (in_substep_) {
            deferred_.push_back({true, BodyId{}, initial, shape});
            return BodyId{};
        }
        pull();
        uint32_t index;
        if (!free_.empty()) {
            index = free_.back();
            free_.pop_back();
        } else {
            index = static_cast<uint32_t>(bodies_.size());
            bodies_.push_back(pxg_body{});
            shapes_.push_back(no_shape());
            generations_.push_back(0);
        }
        pxg_body& b = bodies_[index];
        b = initial;
        b.alive = 1;
        std::memcpy(b.prev_position, b.position, sizeof(b.position));
        std::memcpy(b.prev_orientation, b.orientation, sizeof(b.orientation));
        pxg_shape s = shape;
        s.owner = index;
        shapes_[index] = s;
        new_shapes_.push_back(index);
        structure_dirty_ = true;
        return BodyId{index, generations_[index]};
    }

    void remove_body(BodyId id) {
        if (!is_valid(id)) return;
        if (in_substep_) {
            deferred_.push_back({false, id, pxg_body{}, no_shape()});
            return;
        }
        pull();
        bodies_[id.index].alive = 0;
        shapes_[id.index].enabled = 0;
        ++generations_[id.index];
        free_.push_back(id.index);
        structure_dirty_ = true;
    }

    bool is_valid(BodyId id) const {
        return id.is_valid() && id.index < bodies_.size() && generations_[id.index] == id.generation &&
               alive_on_host(id.index);
    }

    // The body `id` addresses, or nullptr for a stale handle (world.h:347). Writes through the
    // pointer reach the GPU before the next step.
    pxg_body* get_body(BodyId id) {
        if (!is_valid(id)) return nullptr;
        pull();
        bodies_dirty_ = true;
        return &bodies_[id.index];
    }
    const pxg_body* get_body(BodyId id) const {
        if (!is_valid(id)) return nullptr;
        const_cast<PhysicsWorld*>(this)->pull();
        return &bodies_[id.index];
    }
    // The primary shape of `id` (world.h:367-374); edits apply before the next step.
    pxg_shape* get_shape(BodyId id) {
        if (!is_valid(id)) return nullptr;
        structure_dirty_ = true;
        new_shapes_.push_back(id.index);  // its proxy is re-created
        return &shapes_[id.index];
    }

    uint32_t body_count() const {
        uint32_t n = 0;
        for (size_t i = 0; i < bodies_.size(); ++i) n += alive_on_host(static_cast<uint32_t>(i)) ? 1u : 0u;
        return n;
    }

    LayerMatrix& layers() { return layers_; }
    const LayerMatrix& layers() const { return layers_; }

    // Mutable like the reference's config() (world.h:1207); changes apply at the next step.
    pxg_config& config() { return cfg_; }
    const pxg_config& config() const { return cfg_; }

    void set_gravity(float x, float y, float z) {
        cfg_.gravity[0] = x;
        cfg_.gravity[1] = y;
        cfg_.gravity[2] = z;
    }
    const float* gravity() const { return cfg_.gravity; }

    // --- stepping (world.h:1229-1317) ---
    // Accumulates dt (clamped to 0.25 s) and runs up to max_substeps fixed substeps.
    uint32_t step(float dt) {
        const float frame = dt < 0.25f ? dt : 0.25f;
        accumulator_ += frame;
        const float h = cfg_.fixed_dt;
        uint32_t n = 0;
        while (accumulator_ >= h && n < cfg_.max_substeps) {
            step_fixed(h);
            accumulator_ -= h;
            ++n;
        }
        return n;
    }

    void step_fixed(float h) {
        std::vector<Deferred> cmds;
        cmds.swap(deferred_);
        for (const Deferred& c : cmds) {
            if (c.create) add_body(c.body, c.shape);
            else remove_body(c.id);
        }
        push();
        if (on_substep.empty()) {
            check(pxg_step_fixed(w_, h), "step_fixed");
        } else {
            check(pxg_step_phase(w_, PXG_PHASE_FORCES_COLLIDE, h), "step_fixed (collide)");
            state_on_device_ = true;
            bodies_dirty_ = false;
            in_substep_ = true;
            on_substep.emit(*this, h);
            in_substep_ = false;
            if (bodies_dirty_) {
                check(pxg_set_bodies(w_, bodies_.data(), static_cast<uint32_t>(bodies_.size())), "on_substep upload");
                bodies_dirty_ = false;
            }
            check(pxg_step_phase(w_, PXG_PHASE_SOLVE_INTEGRATE, h), "step_fixed (solve)");
        }
        state_on_device_ = true;
        manifolds_valid_ = false;
    }

    // This step's valid manifolds with the solver's final impulses (world.h:731).
    const std::vector<pxg_manifold>& manifolds() const {
        if (!manifolds_valid_) {
            uint32_t n = 0;
            check(pxg_get_manifolds(w_, nullptr, 0, &n), "manifolds");
            manifolds_.resize(n);
            if (n) check(pxg_get_manifolds(w_, manifolds_.data(), n, &n), "manifolds");
            manifolds_.resize(n);
            manifolds_valid_ = true;
        }
        return manifolds_;
    }

    // FNV-1a over alive bodies' pose and velocities (world.h:1334-1359).
    uint64_t world_state_hash() {
        push();
        uint64_t h = 0;
        check(pxg_state_hash(w_, &h), "world_state_hash");
        return h;
    }

    pxg_stats stats() const {
        pxg_stats s;
        check(pxg_get_stats(w_, &s), "stats");
        return s;
    }

    pxg_world* handle() { return w_; }

private:
    // A new handle with room for at least n bodies; the warm-start cache moves over (the caller
    // restores the fat AABBs).
    void grow(uint32_t n) {
        uint32_t cap = cfg_.max_bodies;
        while (cap < n) cap *= 2;
        uint32_t nk = 0;
        check(pxg_get_warm_cache(w_, nullptr, nullptr, 0, &nk), "warm cache");
        std::vector<uint64_t> keys(nk);
        std::vector<float> imp(size_t(nk) * 3);
        if (nk) check(pxg_get_warm_cache(w_, keys.data(), imp.data(), nk, &nk), "warm cache");
        pxg_config c = cfg_;
        c.max_bodies = c.max_shapes = cap;
        pxg_world* nw = nullptr;
        check(pxg_create(&c, &nw), "pxg_create");
        pxg_destroy(w_);
        w_ = nw;
        cfg_ = c;
        applied_cfg_ = c;
        check(pxg_set_layer_masks(w_, layers_.rows()), "layers");
        pending_cache_keys_ = std::move(keys);
        pending_cache_imp_ = std::move(imp);
    }

public:

private:
    bool alive_on_host(uint32_t i) const {
        // aliveness changes only on the host (add/remove), so the mirror's flag is current
        return bodies_[i].alive != 0;
    }

    // device -> host mirror
    void pull() {
        // between the two phases this is the collide phase's state (forces integrated)
        if (!state_on_device_ || bodies_.empty()) return;
        check(pxg_get_bodies(w_, bodies_.data(), static_cast<uint32_t>(bodies_.size())), "get_body");
        state_on_device_ = false;
    }

    // host -> device before a step
    void push() {
        if (std::memcmp(&applied_cfg_, &cfg_, sizeof(cfg_)) != 0) {
            check(pxg_set_config(w_, &cfg_), "config");
            applied_cfg_ = cfg_;
        }
        if (layers_.version() != applied_layers_) {
            check(pxg_set_layer_masks(w_, layers_.rows()), "layers");
            applied_layers_ = layers_.version();
        }
        const uint32_t n = static_cast<uint32_t>(bodies_.size());
        if (structure_dirty_) {
            // the broadphase keeps each surviving shape's fat AABB; new shapes get tight + margin
            std::vector<float> fat(size_t(uploaded_shapes_) * 6);
            if (uploaded_shapes_) check(pxg_get_fat_aabbs(w_, fat.data(), uploaded_shapes_), "fat AABBs");
            if (n > cfg_.max_bodies || n > cfg_.max_shapes) grow(n);
            check(pxg_set_bodies(w_, bodies_.data(), n), "bodies");
            check(pxg_set_shapes(w_, shapes_.data(), n), "shapes");
            std::vector<float> now(size_t(n) * 6);
            check(pxg_get_fat_aabbs(w_, now.data(), n), "fat AABBs");
            std::vector<char> fresh(n, 0);
            for (uint32_t i : new_shapes_) fresh[i] = 1;
            for (uint32_t i = 0; i < uploaded_shapes_ && i < n; ++i)
                if (!fresh[i]) std::memcpy(&now[size_t(i) * 6], &fat[size_t(i) * 6], 6 * sizeof(float));
            check(pxg_set_fat_aabbs(w_, now.data(), n), "fat AABBs");
            if (!pending_cache_keys_.empty()) {
                check(pxg_set_warm_cache(w_, pending_cache_keys_.data(), pending_cache_imp_.data(),
                                         static_cast<uint32_t>(pending_cache_keys_.size())), "warm cache");
                pending_cache_keys_.clear();
                pending_cache_imp_.clear();
            }
            uploaded_shapes_ = n;
            new_shapes_.clear();
            structure_dirty_ = false;
            bodies_dirty_ = false;
        } else if (bodies_dirty_) {
            check(pxg_set_bodies(w_, bodies_.data(), n), "bodies");
            bodies_dirty_ = false;
        }
    }

    struct Deferred {
        bool create;
        BodyId id;
        pxg_body body;
        pxg_shape shape;
    };
    std::vector<Deferred> deferred_;
    pxg_world* w_ = nullptr;
    pxg_config cfg_, applied_cfg_;
    LayerMatrix layers_;
    uint64_t applied_layers_ = 0;
    std::vector<pxg_body> bodies_;
    std::vector<pxg_shape> shapes_;  // shape slot i belongs to body i
    std::vector<uint32_t> generations_, free_, new_shapes_;
    std::vector<uint64_t> pending_cache_keys_;
    std::vector<float> pending_cache_imp_;
    uint32_t uploaded_shapes_ = 0;
    bool structure_dirty_ = false, bodies_dirty_ = false, state_on_device_ = false, in_substep_ = false;
    float accumulator_ = 0.0f;
    mutable std::vector<pxg_manifold> manifolds_;
    mutable bool manifolds_valid_ = false;
};

}  // namespace pxg
