// Device records of the cloths and hair grooms one world steps (pxg_set_cloths /
// pxg_set_strand_sets), shared by the host upload and the kernels in pxg_soft.hip. Every cloth or
// groom owns contiguous ranges of the flat arrays below; its record holds the offsets, counts and
// parameters. One workgroup advances one cloth or groom per frame, its particle positions in LDS.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pxg_state.h"

namespace pxg {

constexpr uint32_t kSoftMaxParticles = 8192;  // per cloth or groom: positions held in LDS (128 KB)
constexpr uint32_t kSoftMaxColliders = 256;   // colliders one cloth or groom meets in a frame
constexpr uint32_t kSoftBlock = 256;
constexpr uint32_t kNoBody = 0xFFFFFFFFu;

struct ClothRec {
    uint32_t p0, np, c0, n_stretch, n_bend, b0, n_sb, n_bb;  // particles, constraints (stretch, bend), batches
    uint32_t t0, n_tether, a0, n_anchor, tr0, n_tri, pt0, columns;  // tethers, anchors, triangles, triangle lists
    uint32_t alive, enabled, self_collision, substeps, iterations, layer, bins0, buckets;
    float stretch_compliance, bend_compliance, damping, thickness, friction, gravity_scale, max_velocity, wind_turbulence;
    float air_drag, air_lift, self_distance, sleep_threshold, sleep_time;
    uint32_t tether_serial;  // a tether's anchor is another tether's particle: tethers run in order on one lane
    uint32_t reserved0, reserved1;
    float ext[4], wind[4];
};

struct StrandRec {
    uint32_t p0, np, c0, n_stretch, b0, n_b, s0, n_strands;  // particles, constraints, batches, strands
    uint32_t cb0, n_cb, alive, enabled, substeps, iterations, layer, reserved;  // collide bodies
    float stretch_compliance, global_shape, global_shape_falloff, local_shape, damping, thickness, friction, gravity_scale;
    float max_velocity, wind_turbulence, air_drag, sleep_threshold, sleep_time, reserved0, reserved1, reserved2;
    float ext[4], wind[4];
};

struct SoftState {  // per cloth or groom, read back by the host
    uint32_t awake;
    float sleep_timer, time;
    uint32_t error;  // 1: more than kSoftMaxColliders colliders met in a frame
    float bmin[4], bmax[4];
};

struct SoftDev {
    // cloths
    const ClothRec* cl;
    SoftState* cl_state;
    uint32_t n_cloths;
    float4 *cl_pos, *cl_prev, *cl_vel;  // pos.w = inv_mass
    uint4* cl_con;                      // a, b, rest length bits, multiplier bits: a cloth's stretch then bend
    const uint2* cl_batch;              // [begin, end) in the cloth's constraints; bend batches after stretch ones
    const uint4* cl_tether;             // particle, anchor, max length bits
    float4* cl_anchor;                  // 3 per anchor: world (xyz, particle bits), prev world, local
    const uint32_t* cl_anchor_body;     // per anchor: the live body it follows, or kNoBody (a static or stale pin)
    const uint4* cl_tri;                // triangle corners
    const uint32_t* cl_ptri_off;        // per cloth np + 1 offsets into cl_ptri (at pt0)
    const uint32_t* cl_ptri;            // per particle: the triangles touching it, ascending, once per corner
    float4* cl_share;                   // per triangle: a third of its aerodynamic force
    uint32_t* cl_bins;                  // self-collision: cell of each particle, bucket ends, particles by bucket
    uint32_t* cl_cols;                  // kSoftMaxColliders collider shape slots per cloth
    // hair
    const StrandRec* st;
    SoftState* st_state;
    uint32_t n_sets;
    float4 *st_pos, *st_prev, *st_vel;
    uint4* st_con;
    const uint2* st_batch;
    const float4* st_rest_local;
    const float4* st_rest_offset;
    float4* st_strand;                  // 4 per strand: root (xyz, first bits), root q, prev root (xyz, count bits), prev root q
    const uint32_t* st_collide;         // collide-body indices, kNoBody for a stale one
    uint32_t* st_cols;
};

// frame: the clamped frame time; frame_substeps / world_iters: the world's cloth substeps for it
// and iterations; max_np: the largest particle count (dynamic LDS of the launch)
void launch_refit(const DevState& d, float h, float margin, hipStream_t st);
void launch_cloths(const DevState& d, const SoftDev& s, uint32_t max_np, float frame, uint32_t frame_substeps,
                   uint32_t world_iters, pxd::v3 gravity, float margin, hipStream_t st);
void launch_strands(const DevState& d, const SoftDev& s, uint32_t max_np, float frame, uint32_t frame_substeps,
                    uint32_t world_iters, pxd::v3 gravity, float margin, hipStream_t st);
hipError_t soft_lds_limit(size_t bytes);  // allows `bytes` of dynamic LDS for both kernels

}  // namespace pxg
