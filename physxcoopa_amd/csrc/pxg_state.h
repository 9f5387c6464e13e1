// Device memory layout of one world and the launchers of its step kernels.
//
// Bodies and shapes are stored structure-of-arrays with 16-byte records so that a wavefront's
// loads of one field coalesce; the solver works on compact per-step copies laid out in solve
// order (colour by colour), so each colour's manifolds are one contiguous range.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pxg_device.h"
#include "../../include/pxg.h"

namespace pxg {

struct MlpDev;
constexpr int kBlock = 256;
// grid barrier words: [0] arrivals, [1] set-up arrivals, [2 + x] blocks on XCD x (near barriers)
// grid barrier words: the counting barrier (0), the XCD pick (1-17), and from kBarFlags one 64-byte
// line per block of a one-XCD launch for the flag form of the barrier (at most kBarFlagBlocks)
constexpr uint32_t kBarFlags = 32, kBarFlagBlocks = 64;
constexpr uint32_t kBarWords = kBarFlags + 16 * kBarFlagBlocks;
// XCDs of an MI355X: a one-XCD solve launches this many blocks per block that carries on
constexpr uint32_t kXcds = 8;
constexpr int kPersistBlock = 512;  // persistent cooperative kernels: 8 waves per block (up to 256 VGPRs each)
constexpr int kMaxColours = 254;     // colours per step (reference order: chain layers)
constexpr int kMaskColours = 64;     // greedy colouring: colours tracked in 64-bit per-body masks
constexpr int kBuckets = kMaxColours + 1;  // colours + the "skipped by the solver" group
// buckets of the solve-order sort: each colour split into its multi-point and its one-point
// manifolds (in that order), then the skipped group
constexpr int kSortBuckets = 2 * kMaxColours + 1;
constexpr int kMaxLarge = 256;       // shapes on the brute-force broadphase path
constexpr uint32_t kMaxCells = 1u << 18;
constexpr int kPairLanes = 9;        // broadphase threads per emitter shape
constexpr uint32_t kMaxShapes = (1u << 20) / kPairLanes;  // the pair-count scan covers 2^20 entries
// tiled solve order (pxg_tiles.hip): tiles (one per CU of one XCD), as x strips times y tiles;
// colours over the phases; the sort bucket of the manifolds the solver skips
constexpr uint32_t kTiles = 32, kTileNx = 8, kTileNy = 4;
constexpr uint32_t kTileColours = 64;
constexpr uint32_t kTileSkipBucket = kTileColours * kTiles;

// body flag bits
constexpr uint32_t BF_TYPE = 3u, BF_ALIVE = 1u << 2, BF_AWAKE = 1u << 3, BF_GRAVITY = 1u << 4;
constexpr uint32_t BF_INTERP_NONE = 1u << 5;  // render interpolation off (carried, never read by the step)
constexpr int BF_CONSTRAINT_SHIFT = 8;
// shape info bits (s_info.y)
constexpr uint32_t SF_ENABLED = 1u << 4, SF_TRIGGER = 1u << 5, SF_STATIC = 1u << 6;

// error bits reported through DevState::err
constexpr uint32_t ERR_PAIRS = 1u, ERR_LARGE = 2u, ERR_COLOURS = 4u, ERR_MANIFOLDS = 8u, ERR_CACHE = 16u,
                   ERR_CELLS = 32u, ERR_MESH_VERTICES = 64u;

struct Counters {  // device-side counters, read back by the host at the step's sync points
    uint32_t pairs;
    uint32_t manifolds;
    uint32_t points;
    uint32_t active;        // manifolds the solver runs
    uint32_t ordered;       // active + skipped non-trigger manifolds (cache rebuild list)
    uint32_t remaining;     // colouring: active manifolds not yet coloured
    uint32_t colours;
    uint32_t n_large;
    uint32_t err;
    uint32_t awake;
    uint32_t predicted;
    uint32_t residual_bits; // float bits of the last velocity residual (max, >= 0)
    uint32_t red[12];       // broadphase reductions (sortable-float encoded)
    uint32_t colour_off[kBuckets + 1];
    float grid[12];         // ox, oy, oz, inv_cell, ext_y, ext_z, nx, ny, nz (as floats)
    unsigned long long prof_vel_bytes;  // algorithmic bytes of profiled velocity passes
    unsigned long long prof_stage[4];   // profiled tiled solves, workgroup 0 (10 ns ticks): colours,
                                        // staging bodies in and out, grid barriers, residuals
    uint32_t passes;        // velocity passes run by the last solve
    uint32_t res_slot[2];   // residual max-reduction slots (float bits), alternating per evaluation
    uint32_t ev_n;          // contact events queued (<= the queue's capacity)
    uint32_t ev_dropped;    // events lost to a full queue since it was last cleared
    uint32_t ev_snap_n[2];  // entries of the two pair snapshots
    uint32_t tile_n;        // tiled order: awake dynamic bodies
    uint32_t tile_cols[3];  // tiled order: colours of each phase
    uint32_t tile_wg;       // tiled solve: workgroups | workgroups that held their tile in LDS << 8
    uint32_t tile_need;     // tiled solve: the most LDS bytes a workgroup's tile needed
    uint32_t tile_nu;       // tiled order: entries of all group tables
    uint32_t tile_ncomp;    // tiled order: components of the crossing manifolds
};

struct DevState {
    uint32_t nb, ns;              // bodies, shapes in use
    uint32_t cap_pairs, cap_manifolds, cache_mask;
    // bodies
    float4 *b_pos;      // position, sleep_timer
    float4 *b_rot;      // x, y, z, w
    float4 *b_vel;      // linear velocity, inv_mass
    float4 *b_angv;     // angular velocity, mass
    float4 *b_force;    // force accumulator, linear drag
    float4 *b_torque;   // torque accumulator, angular drag
    float4 *b_iil;      // inverse inertia (local diagonal), -
    float4 *b_iiw;      // 3 records per body: world inverse inertia, column-major 9 floats + pad
    float4 *b_prev_pos, *b_prev_rot;
    uint32_t *b_flags;
    // shapes
    uint4 *s_info;      // owner, type | flags | caxis << 8, layer, fcomb | rcomb << 8
    float4 *s_lc;       // local centre, sphere radius
    float4 *s_lr;       // local rotation
    float4 *s_dim;      // box half extents, capsule radius
    float4 *s_mat;      // capsule half height, dynamic friction, restitution, static friction
    // convex hulls (pxg_set_hulls): the table PXG_CONVEX_HULL shapes index (pxg_hull.h's HullTab)
    uint4 *h_meta, *h_meta2;
    float4 *h_lb, *h_v, *h_pl, *h_cen;
    uint2 *h_face, *h_e;
    uint32_t *h_fv;
    uint32_t n_hulls;
    // triangle meshes (pxg_set_meshes): the table PXG_TRIANGLE_MESH shapes index (pxg_mesh.h's MeshTab)
    uint4 *me_meta, *me_tri, *me_adj;
    float4 *me_lb, *me_v, *me_n, *me_node_lo, *me_node_hi;
    uint2 *me_node_tri;
    uint32_t *me_order;
    uint32_t n_meshes;
    float4 *s_fmin, *s_fmax;   // fat AABB (persistent broadphase state)
    float4 *s_tmin, *s_tmax;   // tight AABB of this step; w of s_tmin: 1 = large shape
    uint32_t *layer_mask;      // 32 rows
    // broadphase
    uint32_t *cell_count, *cell_start, *cell_cursor;
    float4 *ci_fmin, *ci_fmax, *ci_tmin, *ci_tmax;  // cell-ordered copies: fat/tight boxes, w = index /
                                                    // owner / static flag bits
    uint32_t *large_list;
    uint32_t *pair_count, *pair_off;   // per (emitter shape, lane): kPairLanes entries per shape
    uint2 *pairs;                      // (shape a, shape b), a < b
    // pair-indexed narrowphase output
    uint32_t *pm_valid, *pm_off;
    uint2 *pm_ab;
    float4 *pm_nf;
    uint4 *pm_meta;
    float4 *pm_p, *pm_q;
    uint32_t *scan_tmp;   // >= 2048 words of scan scratch
    // manifolds, discovery order (dense)
    uint2 *m_ab;        // body a, body b
    uint2 *m_shapes;    // shape a, shape b
    float4 *m_nf;       // normal, friction
    uint4 *m_meta;      // count | trigger << 8, restitution bits, active, -
    float4 *m_p;        // 4 per manifold: position, penetration
    float4 *m_q;        // 4 per manifold: separation, feature id bits, -, -
    // colouring
    uint32_t *m_colour;
    unsigned long long *col_best;   // per body
    unsigned long long *col_mask;   // per body
    unsigned long long *col_dup;    // per body: colours claimed twice by last step's colours
    // reference solve order: each awake dynamic body's manifolds in pair order
    uint32_t *adj_cnt, *adj_start, *adj_cur;  // per body
    uint32_t *adj;                  // 2 per manifold: the bodies' chains, concatenated
    uint32_t *m_pred;               // 2 per manifold: predecessor on body a / body b (kNoPred: none)
    uint32_t *col_table, *col_table_off;
    uint32_t *order;                // solve position -> manifold index
    // residual mode: the active solve positions split by whether a tail colour (see the solve
    // kernel) writes one of their bodies: those it does not, from the front of ready_list (their
    // count in ready_cnt[0]), the others from the back (ready_cnt[1])
    uint32_t *body_last;            // per body: last colour writing it
    uint32_t *ready_cnt;            // 2 counters
    uint32_t *ready_list;
    // solver arrays, solve order
    uint2 *sm_ab;
    float4 *sm_nf;
    uint32_t *sm_cnt;
    float4 *sp_pos;     // position, penetration
    float4 *sp_aux;     // separation, restitution bias, feature id bits, miss flag
    float4 *sp_imp;     // normal, tangent1, tangent2 impulses
    float4 *sp_row;     // normal, tangent1, tangent2 effective masses; velocity bias (-separation / h, or 0)
    // solver bodies
    float4 *sb_v;       // linear velocity, effective inverse mass
    float4 *sb_w;       // angular velocity, active (1/0)
    float4 *sb_I;       // 3 per body: effective world inverse inertia (9) + position (3)
    float4 *sb_x;       // position pass: position, effective inverse mass
    float4 *sb_ref;     // position pass reference positions
    // tiled solve order (pxg_tiles.hip; allocated on the first tiled step)
    uint32_t *t_body;     // 2 * nb: each awake dynamic body's tile, then each body's crossing-component root
    uint32_t *t_par;      // nb: union-find parents of the crossing components
    uint32_t *t_csz;      // nb: crossing manifolds per component root, then each root's group
    uint32_t *m_tile;     // per manifold: phase << 8 | group
    uint32_t *t_off;      // kTileSkipBucket + 2: first solve position of each (global colour * kTiles
                          // + group); then of the skipped group; then the end of the order
    uint32_t *ts_a;       // 2 per manifold: its bodies' table keys (phase * kTiles + group) << 20 | body
    uint32_t *tf, *tu;    // the sorted keys' first-occurrence flags and their scan
    uint32_t *tb_key, *tb_list;  // per (phase, group), ascending: the bodies its manifolds touch (key, body)
    uint32_t *tb_off;     // 3 * kTiles + 1 offsets into tb_list
    uint32_t *sm_loc;     // per solve position: body a's | body b's index << 16 in the group's table
    unsigned long long *t_col_best, *t_col_mask;  // 3 * nb: the colouring's per-phase body words
    uint32_t *tk_a, *tk_b, *tv_a, *tv_b;         // sort keys and values
    unsigned long long *tk64_a, *tk64_b;
    void *t_sort_tmp;
    size_t t_sort_bytes;
    // warm-start cache: two open-addressing tables, swapped every step
    unsigned long long *ck_keys[2];
    unsigned long long *ck_ord[2];  // last writer in the reference's manifold order
    float4 *ck_val[2];
    int cur;            // table written this step (the other is read)
    // islands
    uint32_t *uf_parent;
    uint32_t *uf_root;  // each body's island root, after the unions
    uint32_t *isl_min;
    // joints (dynamics/joint.h): definitions and accumulators (pxg_joints.h layout), the solve
    // order as layers (no two joints of a layer share a dynamic body), forward then reverse
    uint32_t nj, nj_layers;
    uint4 *j_meta;        // type, body a, body b, live (valid and enabled)
    float4 *j_def;        // 6 per joint
    float4 *j_acc;        // 2 per joint
    uint32_t *j_order;    // 2 * nj joint indices: forward layering, then reverse layering
    uint32_t *j_layer_off;// 2 * (nj_layers + 1) offsets into j_order's two halves
    // body pairs whose contacts are off (jointed without collide_connected, or ignored),
    // sorted keys lo << 32 | hi
    unsigned long long *nc_keys;
    uint32_t n_nc;
    uint32_t *bar;      // grid barrier of the persistent kernels: arrival count (zeroed per launch)
    Counters *cnt;
    // learned warm start
    const struct MlpDev *mlp;  // BF16 network (nullptr = off)
    float *feat;        // learned: compact feature rows of misses; recording: rows per point (4 s + i)
    uint32_t *miss_idx; // learned: point index of each compact row
    float *pred;        // learned: clamped predictions, 3 per compact row
    uint32_t record;    // 1: write every active point's features (dataset recording)
    // contact events: each step's confirmed shape pairs (its valid manifolds) as sorted keys
    // kind << 34 | shape a << 17 | shape b (kind 0 solid, 1 trigger), double-buffered (this
    // step's in ev_snap[cur], last step's in the other, counts in Counters::ev_snap_n); the
    // unsorted keys and the sort's scratch; per snapshot entry of both steps its event category,
    // per block of 1024 entries and category the event count, then the queue position; the queue
    unsigned long long *ev_keys, *ev_snap[2];  // cap_pairs each
    void *ev_sort_tmp;
    size_t ev_sort_bytes;
    uint8_t *ev_code;   // 2 * cap_pairs: this step's entries, then last step's
    uint32_t *ev_boff;  // 6 categories x (2 * ceil(cap_pairs / 1024) + 2) blocks
    uint4 *ev_out;      // body a, body b, type (0 enter, 1 stay, 2 exit), is_trigger
    uint32_t cap_events;
};

constexpr uint32_t kNoPred = 0xFFFFFFFFu;

struct StepParams {
    float h;
    float gx, gy, gz;
    float margin, large_extent;
    float pos_corr, slop, max_corr, rest_thresh;
    float sleep_lin, sleep_ang, sleep_time, max_lin_vel;
    uint32_t exact_quat;
    uint32_t profile;   // 1: velocity passes count their algorithmic bytes
};

// launchers (pxg_kernels.hip); all asynchronous on `st`. Where a launcher takes a count bound,
// the grid covers the bound and the kernels read the step's actual count from Counters.
void launch_integrate_forces(const DevState& d, const StepParams& p, hipStream_t st);
void launch_broadphase_count(const DevState& d, const StepParams& p, hipStream_t st);
void launch_broadphase_write(const DevState& d, const StepParams& p, hipStream_t st);
void launch_narrowphase(const DevState& d, const StepParams& p, uint32_t n_pairs, hipStream_t st);
void launch_colour_init(const DevState& d, uint32_t n_manifolds, hipStream_t st);
void launch_colour_rounds(const DevState& d, uint32_t n_manifolds, uint32_t first_round, uint32_t rounds, hipStream_t st);
size_t col_table_words(size_t max_blocks);  // words of col_table for up to max_blocks manifold blocks
void launch_build_order(const DevState& d, uint32_t n_manifolds, hipStream_t st);
void launch_solver_setup(const DevState& d, const StepParams& p, uint32_t bound, hipStream_t st);
void launch_warm_apply(const DevState& d, uint32_t begin, uint32_t end, hipStream_t st);
void launch_velocity(const DevState& d, const StepParams& p, uint32_t begin, uint32_t end, int use_rest, uint32_t offset,
                     hipStream_t st);
void launch_residual(const DevState& d, const StepParams& p, uint32_t n_active, int use_rest, hipStream_t st);
void launch_position_setup(const DevState& d, hipStream_t st);
void launch_set_passes(const DevState& d, uint32_t passes, hipStream_t st);  // residual mode, eager path
void launch_position(const DevState& d, const StepParams& p, uint32_t begin, uint32_t end, hipStream_t st);
void launch_solver_finish(const DevState& d, uint32_t bound, hipStream_t st);
void launch_islands_sleep(const DevState& d, const StepParams& p, uint32_t bound, hipStream_t st);
void launch_integrate_velocities(const DevState& d, const StepParams& p, hipStream_t st);
// joints: wake_jointed_islands before the solve; the eager path's per-layer launches (layer
// [begin, end) of the forward order, or of the reverse order when rev); the post-integrate
// position passes and world inertia update (world.h:1310-1314)
void launch_wake_jointed(const DevState& d, hipStream_t st);
void launch_joint_warm(const DevState& d, uint32_t begin, uint32_t end, hipStream_t st);
void launch_joint_velocity(const DevState& d, const StepParams& p, int rev, uint32_t begin, uint32_t end, hipStream_t st);
void launch_joint_position(const DevState& d, const StepParams& p, uint32_t begin, uint32_t end, hipStream_t st);
void launch_joints_post(const DevState& d, const StepParams& p, uint32_t pos_iters, hipStream_t st);
void launch_init_bodies(const DevState& d, hipStream_t st);
// staged records (9 float4 fields, SoA over n) and flags into bodies idx[0..n); `inertia`: recompute
// their world inverse inertia
void launch_scatter_bodies(const DevState& d, const uint32_t* idx, const float4* rec, const uint32_t* flags, uint32_t n,
                           int inertia, hipStream_t st);
// contact events of this step (world.h:1964-2018, pxg_events.hip): sorts this step's valid
// manifolds (at most n_pairs) into the snapshot ev_snap[cur], compares it with last step's (at
// most prev_bound entries) and appends the events to the queue
size_t events_sort_bytes(uint32_t cap_pairs);
// scene raycasts (pxg_query.hip). Per batch: per-shape world geometry (3 float4 per shape),
// query info (uint4) and the bounds of the live fat-box centres (6 order-encoded floats, reset
// to empty before the prep) from the current poses; then shapes and rays sorted by Morton code,
// the sorted shape records and their tiles' boxes, and the closest (or any) hit of n rays (2
// float4 each; hits 3 float4 each). raycast_all: every hit of one ray (unordered, at most cap,
// count in *count) from the unsorted records.
struct QueryFilterDev {
    uint32_t layer_mask, include_triggers, ignore_body;
};
struct QueryScratch {
    float4* geom;       // 3 per shape slot
    uint4* info;        // per shape slot
    uint32_t* bounds6;
    uint32_t *skey, *skey2, *sidx, *sidx2;  // per shape slot
    uint32_t *rkey, *rkey2, *ridx, *ridx2;  // per ray
    void* sorted;       // per shape slot: query_shape_record_bytes() each
    float4* chunk_box;  // box of the live fat AABBs of each 64-slot chunk, in slot order: 2 per chunk
    float4* tile_box;   // 2 per tile of 256 shapes
    void* sort_tmp;
    size_t sort_bytes;
};
// overlap query shape: type (SH_*), sphere (c, r), box (c, he, q), capsule (a, b, r)
struct OverlapDev {
    uint32_t type;
    float r;
    float4 c, he, q, a, b;
};
void launch_overlap(const DevState& d, const float4* geom, const uint4* qinfo, const OverlapDev& q, const QueryFilterDev& f,
                    uint32_t* flags, hipStream_t st);
// shape cast: the cast shape (sphere or capsule, decoded), its pose, direction and max distance
struct CastDev {
    pxd::ShapeGeom shape;
    float4 pos, rot, dir;
    float maxd;
};
// per-shape records (3 float4 per shape), the nearest key (reset to ~0 before), the hit (3 float4)
void launch_shape_cast(const DevState& d, const uint4* qinfo, const CastDev& c, const QueryFilterDev& f, float4* rec,
                       unsigned long long* best, float4* hit, hipStream_t st);
// compute_penetration: the posed probe in c (dir and maxd unused) against every shape; each
// penetration appends (normal, depth; point, slot; owner) as 3 float4 at rec[3 * atomicAdd(count, 1)]
// (count reset to 0 before; room for 3 float4 per shape)
void launch_penetration(const DevState& d, const uint4* qinfo, const CastDev& c, const QueryFilterDev& f, float4* rec,
                        uint32_t* count, hipStream_t st);
// many shape casts at once: the shapes binned into Morton-ordered tiles, then one wavefront per
// cast over the tiles its swept AABB meets; hits: 3 float4 per cast (the pxg_raycast_hit layout)
hipError_t launch_shape_cast_batch(const DevState& d, const QueryScratch& q, const CastDev* casts, uint32_t n,
                                   const QueryFilterDev& f, float4* hits, hipStream_t st);
// many overlap queries, one wavefront each: counts[i] overlapping shapes, the owners of the first
// min(cap, counts[i]) in slot order at bodies[i * cap]; chunk_box (2 float4 per 64-slot chunk) is
// filled first and lets a wavefront skip the chunks its query's bounds miss
void launch_overlap_batch(const DevState& d, const float4* geom, const uint4* qinfo, float4* chunk_box,
                          const OverlapDev* queries, uint32_t n, const QueryFilterDev& f, uint32_t* bodies, uint32_t cap,
                          uint32_t* counts, hipStream_t st);
// the same over ABI query records in device memory (an unknown shape type: count PXG_NO_BODY)
void launch_overlap_batch_records(const DevState& d, const float4* geom, const uint4* qinfo, float4* chunk_box,
                                  const pxg_overlap_query* queries, uint32_t n, const QueryFilterDev& f, uint32_t* bodies,
                                  uint32_t cap, uint32_t* counts, hipStream_t st);
// many penetration probes (dir and maxd unused), one wavefront each, chunk-culled like the batched
// overlaps: counts[i] penetrations, the first min(cap, counts[i]) in slot order as pxg_penetration
// records (10 words) at out[i * cap * 10]
void launch_penetration_batch(const DevState& d, const uint4* qinfo, float4* chunk_box, const CastDev* probes, uint32_t n,
                              const QueryFilterDev& f, float* out, uint32_t cap, uint32_t* counts, hipStream_t st);
size_t query_sort_bytes(uint32_t cap);
size_t query_shape_record_bytes();
void launch_query_prep(const DevState& d, float4* geom, uint4* qinfo, uint32_t* bounds6, hipStream_t st);
hipError_t launch_raycast(const DevState& d, const QueryScratch& q, const float4* rays, uint32_t n, const QueryFilterDev& f,
                          uint32_t any, float4* hits, hipStream_t st);
void launch_raycast_all(const DevState& d, const float4* geom, const uint4* qinfo, const float4* ray,
                        const QueryFilterDev& f, float4* hits, uint32_t cap, uint32_t* count, hipStream_t st);
hipError_t launch_events(const DevState& d, uint32_t n_pairs, uint32_t prev_bound, int cur, hipStream_t st);
void launch_apply_predictions(const DevState& d, uint32_t n_bound, hipStream_t st);
void launch_init_proxies(const DevState& d, const StepParams& p, hipStream_t st);
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st);
void launch_fill_u64(unsigned long long* p, unsigned long long v, size_t n, hipStream_t st);
void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, uint32_t* scratch, hipStream_t st);
// persistent cooperative kernels (kPersistBlock threads per block, grid barrier on d.bar);
// cfgv = vel, relax, pos iterations, residual mode, residual max passes, one XCD, ready split (the
// tiled solve: tiles in LDS allowed), residual check interval
// algorithmic bytes of the last persistent solve into cnt->prof_vel_bytes (profiling)
void launch_solve_bytes(const DevState& d, const uint32_t cfgv[8], uint32_t bound, hipStream_t st);
// splits the active solve positions for the residual computed during the tail colours
void launch_ready_split(const DevState& d, uint32_t bound, hipStream_t st);
hipError_t launch_solve_persistent(const DevState& d, const StepParams& p, const uint32_t cfgv[8], float tol,
                                   uint32_t nblocks, hipStream_t st);
// colours every active manifold: last step's colours where still valid, then Jones-Plassmann
// rounds (at most max_rounds) for the rest
hipError_t launch_colour_persistent(const DevState& d, uint32_t max_rounds, uint32_t nblocks, uint32_t reference,
                                    hipStream_t st);
// the reference-order colouring with one launch per stage (rounds launched by the host)
void launch_reference_chains(const DevState& d, uint32_t n_manifolds, hipStream_t st);
void launch_reference_layers(const DevState& d, uint32_t n_manifolds, uint32_t rounds, hipStream_t st);
int max_coresident_blocks(int device);
// tiled solve order (pxg_tiles.hip): sort scratch bytes; the tiles and each active manifold's
// phase and group (after launch_colour_init, before the colouring); the order, colour table and
// body tables (after the colouring); each solve position's body indices (after the solver set-up)
size_t tile_sort_bytes(uint32_t cap_bodies, uint32_t cap_items);
hipError_t launch_tile_setup(const DevState& d, uint32_t n_bound, hipStream_t st);
hipError_t launch_tile_order(const DevState& d, uint32_t n_bound, hipStream_t st);
void launch_tile_slots(const DevState& d, uint32_t n_bound, hipStream_t st);
// the tiled solve: warm start, velocity (fixed or residual mode), relax and position passes of
// the tiled order, each tile's colours in one workgroup (its bodies and impulses in LDS)
hipError_t launch_solve_tiled(const DevState& d, const StepParams& p, const uint32_t cfgv[8], float tol, uint32_t nblocks,
                              hipStream_t st);
// a grid barrier's bounded wait expired (its own bit: the others above report capacity limits)
constexpr uint32_t ERR_BARRIER_BIT = 128u;

}  // namespace pxg
