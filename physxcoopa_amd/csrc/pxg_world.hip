// C ABI of the MI355X contact step (include/pxg.h): world handles, device buffers, uploads,
// downloads and the host-side orchestration of one fixed step. The step runs on the handle's own
// HIP stream; the host waits on the device twice per step: once for the pair count (sizes the
// narrowphase) and once for the colour table (sizes the per-colour solver launches).
#include "../../include/pxg.h"
#include "pxg_state.h"
#include "pxg_hostio.h"
#include "pxg_mesh_cook.h"
#include "pxg_mlp.h"
#include "pxg_soft.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace pxg;

static thread_local std::string g_last_error;

static int fail(int code, const std::string& msg) {
    g_last_error = "[pxg] " + msg;
    return code;
}
#define HIPCHK(call)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) return fail(PXG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct pxg_world {
    pxg_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    DevState d{};
    Counters* h_cnt = nullptr;  // pinned
    std::vector<void*> allocs;
    uint32_t cap_b = 0, cap_s = 0, cap_p = 0, cap_m = 0;
    uint32_t n_bodies = 0, n_shapes = 0;
    std::vector<uint32_t> body_types;
    uint32_t layer_masks[32];
    float accumulator = 0.0f;
    int phase = 0;              // 0: next call collides, 1: collision done, solve pending
    uint32_t last_pairs = 0, last_manifolds = 0, last_ordered = 0, last_active = 0, last_colours = 0;
    bool results_pending = false;  // the last step's counters are still on the device
    uint32_t colour_blocks = 64, solve_blocks = 32;  // caps on the persistent kernels' grids
    bool solve_one_xcd = true;  // the solve runs on the blocks of one XCD with near barriers
    bool tiled_lds = true;      // tiled order: tiles in LDS where they fit (PXG_TILED_LDS=0: global memory)
    // coloured order: colour every step afresh (-1: in residual mode only, where the solve dominates
    // the step) instead of starting from last step's colours
    int colour_fresh = -1;
    uint32_t colour_off[kBuckets + 1];
    pxg_stats stats{};
    uint32_t colour_round = 0;
    bool shapes_ready = false;
    void* noalloc_scratch = nullptr;  // this device's address of `scratch` (pxg_overlap_batch_noalloc)
    uint8_t* body_stage = nullptr;    // staging of pxg_set_bodies_subset (grown on demand)
    uint32_t hull_refs = 0, mesh_refs = 0;  // 1 + the largest hull / mesh index the shapes use
    size_t body_stage_bytes = 0;
    // profiling: HIP event pairs around every velocity-pass launch
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
    double prof_vel_ms = 0.0;
    uint64_t prof_vel_launches = 0;
    int coop_blocks = 0;        // co-resident blocks for the persistent kernels (0: not available)
    MlpDev* d_mlp = nullptr;
    bool eager = false;         // PXG_EAGER=1: one launch per colour instead of persistent kernels
    std::string coop_note;      // why the persistent kernels were dropped, if they were
    // joints: the uploaded table (accumulators refreshed on download), the layer offsets of the
    // forward and reverse orders, and the ignored body pairs (sorted keys)
    std::vector<pxg_joint> joints;
    std::vector<uint32_t> jl_fwd, jl_rev;
    std::vector<uint64_t> ignored;
    uint32_t cap_j = 0, cap_nc = 0;
    // contact events: on by default (the reference always classifies, world.h:1977-1978); the
    // snapshot written this step, last step's pair count, the queue capacity (0: default)
    bool ev_on = true;
    int ev_cur = 0;
    uint32_t ev_prev_n = 0;
    uint32_t ev_cap_req = 0;
    // scene queries: per-shape records and sort scratch, rays and hits (grown on demand)
    QueryScratch q{};
    float4* q_rays = nullptr;
    float4* q_hits = nullptr;
    uint32_t* q_count = nullptr;
    uint32_t q_cap = 0;
    unsigned long long* q_best = nullptr;  // shape casts: nearest (distance, slot) key
    CastDev* q_casts = nullptr;  // batched shape casts, q_casts_cap records
    uint32_t q_casts_cap = 0;
    OverlapDev* q_ov = nullptr;  // batched overlaps: queries and counts (q_ov_cap), bodies (q_ov_out_cap)
    uint32_t* q_ov_counts = nullptr;
    uint32_t* q_ov_out = nullptr;
    uint32_t q_ov_cap = 0;
    size_t q_ov_out_cap = 0;
    float* q_pen = nullptr;  // batched penetrations: q_pen_cap words
    size_t q_pen_cap = 0;
    // bodies' liveness and generations as last uploaded (the device keeps neither generation nor
    // history): anchors and collide bodies resolve against them
    std::vector<uint32_t> body_gen;
    std::vector<uint8_t> body_alive;
    // cloth and hair: device tables (replaced by each upload), host copies of the records, the
    // anchors' and collide bodies' (index, generation) ids, largest particle counts
    SoftDev sd{};
    std::vector<void*> cloth_bufs, strand_bufs;
    std::vector<ClothRec> cl_rec;
    std::vector<StrandRec> st_rec;
    std::vector<uint32_t> cl_anchor_ids;   // 3 per anchor: has_body, body, generation
    std::vector<uint32_t> st_collide_ids;  // 2 per collide body: index, generation
    std::vector<uint32_t> cl_anchor_res, st_collide_res;
    uint32_t cl_max_np = 0, st_max_np = 0;
    bool soft_resolve = false;  // bodies changed since the ids were last resolved
    size_t soft_lds = 0;        // dynamic LDS the soft kernels are allowed
};

extern "C" void pxg_config_default(pxg_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->velocity_iterations = 16;
    c->relax_iterations = 2;
    c->position_iterations = 4;
    c->position_correction = 0.2f;
    c->linear_slop = 0.005f;
    c->max_linear_correction = 0.2f;
    c->restitution_threshold = 0.5f;
    c->sleep_linear = 0.01f;
    c->sleep_angular = 0.02f;
    c->sleep_time = 0.5f;
    c->aabb_margin = 0.1f;
    c->max_substeps = 8;
    c->fixed_dt = 1.0f / 60.0f;
    c->max_linear_velocity = 200.0f;
    c->gravity[2] = -9.81f;
    c->residual_tolerance = 1e-4f;
    c->residual_max_iterations = 256;
    c->residual_check_interval = 1;
    c->large_extent = 16.0f;
    c->cloth_substeps = 4;
    c->cloth_iterations = 1;
}

extern "C" const char* pxg_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* pxg_version(void) { return "pxg 0.1.0 gfx950"; }

template <class T>
static int dalloc(pxg_world* w, T*& p, size_t count) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(PXG_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    w->allocs.push_back(q);
    p = static_cast<T*>(q);
    return PXG_OK;
}
#define DALLOC(ptr, n)                        \
    do {                                      \
        int r_ = dalloc(w, (ptr), (n));       \
        if (r_ != PXG_OK) return r_;          \
    } while (0)

static void free_all(pxg_world* w) {
    for (void* p : w->allocs) (void)hipFree(p);
    w->allocs.clear();
}
// frees one buffer of the world (for buffers that grow)
static void dfree(pxg_world* w, void* p) {
    if (!p) return;
    auto it = std::find(w->allocs.begin(), w->allocs.end(), p);
    if (it != w->allocs.end()) w->allocs.erase(it);
    (void)hipFree(p);
}

static int rebuild_joints(pxg_world* w);
static int step_soft(pxg_world* w, float frame, uint32_t substeps_run);

static StepParams params(const pxg_world* w, float h) {
    StepParams p;
    const pxg_config& c = w->cfg;
    p.h = h;
    p.gx = c.gravity[0]; p.gy = c.gravity[1]; p.gz = c.gravity[2];
    p.margin = c.aabb_margin;
    p.large_extent = c.large_extent > 0.0f ? c.large_extent : 16.0f;
    p.pos_corr = c.position_correction;
    p.slop = c.linear_slop;
    p.max_corr = c.max_linear_correction;
    p.rest_thresh = c.restitution_threshold;
    p.sleep_lin = c.sleep_linear;
    p.sleep_ang = c.sleep_angular;
    p.sleep_time = c.sleep_time;
    p.max_lin_vel = c.max_linear_velocity;
    p.exact_quat = c.use_exact_quaternion_integration;
    p.profile = w->profiling ? 1u : 0u;
    return p;
}

static hipEvent_t take_event(pxg_world* w) {
    if (!w->ev_pool.empty()) {
        hipEvent_t e = w->ev_pool.back();
        w->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// one velocity-pass launch (one colour), bracketed by events when profiling
static void velocity_launch(pxg_world* w, const StepParams& p, uint32_t b, uint32_t e, int rest, uint32_t off) {
    if (e <= b) return;
    if (!w->profiling) {
        launch_velocity(w->d, p, b, e, rest, off, w->stream);
        return;
    }
    hipEvent_t e0 = take_event(w), e1 = take_event(w);
    (void)hipEventRecord(e0, w->stream);
    launch_velocity(w->d, p, b, e, rest, off, w->stream);
    (void)hipEventRecord(e1, w->stream);
    w->ev_pending.push_back({e0, e1});
}

// ---- the scratch of pxg_overlap_batch_noalloc ----
// A file-scope device array, so that entry point needs no allocation: the per-shape world geometry
// (3 float4) and query record (uint4), the scene bounds the geometry pass also writes (6 words) and
// the 64-slot chunk boxes (2 float4), for at most kNoAllocShapes shape slots (about 4 MB, held by
// every device the library runs on until the process ends). In host code `scratch` is not a device
// address: its address on a device comes from hipGetSymbolAddress with that device current.
constexpr uint32_t kNoAllocShapes = 65536u;
constexpr size_t kScrGeom = 0;
constexpr size_t kScrInfo = kScrGeom + size_t(3) * kNoAllocShapes * sizeof(float4);
constexpr size_t kScrBounds = kScrInfo + size_t(kNoAllocShapes) * sizeof(uint4);
constexpr size_t kScrChunks = kScrBounds + 256;
constexpr size_t kScrBytes = kScrChunks + size_t(2) * (kNoAllocShapes / 64u) * sizeof(float4);
static __device__ __align__(256) unsigned char scratch[kScrBytes];

// the address of `scratch` on `device` (current), looked up once per device
static void* noalloc_scratch_address(int device) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        addr[device] = p;
    }
    return addr[device];
}

extern "C" int pxg_create(const pxg_config* cfg, pxg_world** out) {
    if (!out) return fail(PXG_ERR_INVALID, "pxg_create: null output");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PXG_ERR_HIP, "pxg_create: no HIP device");
    pxg_world* w = new pxg_world();
    if (cfg) w->cfg = *cfg;
    else pxg_config_default(&w->cfg);
    for (int i = 0; i < 32; ++i) w->layer_masks[i] = 0xFFFFFFFFu;
    (void)hipGetDevice(&w->device);
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
        delete w;
        return fail(PXG_ERR_HIP, "pxg_create: hipStreamCreate failed");
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&w->h_cnt), sizeof(Counters), hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(w->stream);
        delete w;
        return fail(PXG_ERR_HIP, "pxg_create: hipHostMalloc failed");
    }
    std::memset(w->h_cnt, 0, sizeof(Counters));
    w->noalloc_scratch = noalloc_scratch_address(w->device);  // set-up, once per device
    int coop = 0;
    (void)hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, w->device);
    w->coop_blocks = coop ? max_coresident_blocks(w->device) : 0;
    const char* eager = std::getenv("PXG_EAGER");
    w->eager = eager && eager[0] == '1';
    // grid caps of the persistent kernels: fewer blocks make each grid barrier cheaper
    if (const char* v = std::getenv("PXG_SOLVE_BLOCKS")) w->solve_blocks = std::max(1, std::atoi(v));
    if (const char* v = std::getenv("PXG_COLOUR_BLOCKS")) w->colour_blocks = std::max(1, std::atoi(v));
    if (const char* v = std::getenv("PXG_SOLVE_ONE_XCD")) w->solve_one_xcd = v[0] != '0';
    if (const char* v = std::getenv("PXG_TILED_LDS")) w->tiled_lds = v[0] != '0';
    if (const char* v = std::getenv("PXG_COLOUR_FRESH")) w->colour_fresh = v[0] == '1' ? 1 : 0;
    *out = w;
    return PXG_OK;
}

extern "C" int pxg_destroy(pxg_world* w) {
    if (!w) return fail(PXG_ERR_INVALID, "pxg_destroy: null handle");
    (void)hipStreamSynchronize(w->stream);
    for (auto& pr : w->ev_pending) { w->ev_pool.push_back(pr.first); w->ev_pool.push_back(pr.second); }
    for (hipEvent_t e : w->ev_pool) (void)hipEventDestroy(e);
    free_all(w);
    if (w->h_cnt) (void)hipHostFree(w->h_cnt);
    (void)hipStreamDestroy(w->stream);
    delete w;
    return PXG_OK;
}

extern "C" int pxg_get_config(pxg_world* w, pxg_config* out) {
    if (!w || !out) return fail(PXG_ERR_INVALID, "pxg_get_config: null argument");
    *out = w->cfg;
    return PXG_OK;
}
extern "C" int pxg_set_config(pxg_world* w, const pxg_config* cfg) {
    if (!w || !cfg) return fail(PXG_ERR_INVALID, "pxg_set_config: null argument");
    pxg_config c = *cfg;
    c.max_bodies = w->cfg.max_bodies;
    c.max_shapes = w->cfg.max_shapes;
    c.max_pairs = w->cfg.max_pairs;
    c.max_manifolds = w->cfg.max_manifolds;
    w->cfg = c;
    return PXG_OK;
}
extern "C" int pxg_set_gravity(pxg_world* w, const float g[3]) {
    if (!w || !g) return fail(PXG_ERR_INVALID, "pxg_set_gravity: null argument");
    for (int i = 0; i < 3; ++i) w->cfg.gravity[i] = g[i];
    return PXG_OK;
}

// ---------------------------------------------------------------------------------------------
// allocation
// ---------------------------------------------------------------------------------------------
static int alloc_bodies(pxg_world* w, uint32_t n) {
    uint32_t cap = std::max(n, w->cfg.max_bodies);
    if (w->cap_b) {
        if (n > w->cap_b) return fail(PXG_ERR_CAPACITY, "pxg_set_bodies: more bodies than the world's capacity");
        return PXG_OK;
    }
    w->cap_b = cap;
    DevState& d = w->d;
    DALLOC(d.b_pos, cap); DALLOC(d.b_rot, cap); DALLOC(d.b_vel, cap); DALLOC(d.b_angv, cap);
    DALLOC(d.b_force, cap); DALLOC(d.b_torque, cap); DALLOC(d.b_iil, cap); DALLOC(d.b_iiw, 3 * size_t(cap));
    DALLOC(d.b_prev_pos, cap); DALLOC(d.b_prev_rot, cap); DALLOC(d.b_flags, cap);
    DALLOC(d.col_best, cap); DALLOC(d.col_mask, cap); DALLOC(d.col_dup, cap);
    DALLOC(d.adj_cnt, cap); DALLOC(d.adj_start, cap); DALLOC(d.adj_cur, cap);
    DALLOC(d.sb_v, cap); DALLOC(d.sb_w, cap); DALLOC(d.sb_I, 3 * size_t(cap)); DALLOC(d.sb_x, cap); DALLOC(d.sb_ref, cap);
    DALLOC(d.uf_parent, cap); DALLOC(d.uf_root, cap); DALLOC(d.isl_min, cap); DALLOC(d.body_last, cap);
    DALLOC(d.cnt, 1);
    DALLOC(d.layer_mask, 32);
    DALLOC(d.scan_tmp, 4096);
    DALLOC(d.bar, kBarWords);
    HIPCHK(hipMemsetAsync(d.bar, 0, kBarWords * sizeof(uint32_t), w->stream));
    HIPCHK(hipMemsetAsync(d.cnt, 0, sizeof(Counters), w->stream));
    return PXG_OK;
}

// the tiled solve order's buffers, on the first step that uses it
static int ensure_tiled(pxg_world* w) {
    DevState& d = w->d;
    if (d.t_body) return PXG_OK;
    const size_t B = w->cap_b, M = w->cap_m, P = w->cap_p;
    if (B >= (1u << 20) || M > (1u << 20) || 2 * P > (1u << 22))
        return fail(PXG_ERR_CAPACITY, "tiled solve order: at most 2^20 - 1 bodies, 2^20 manifolds and 2^21 pairs");
    DALLOC(d.t_body, 2 * B); DALLOC(d.t_par, B); DALLOC(d.t_csz, B); DALLOC(d.m_tile, M);
    DALLOC(d.t_off, size_t(kTileSkipBucket) + 16); DALLOC(d.tb_off, 3 * kTiles + 16);
    DALLOC(d.ts_a, 2 * P); DALLOC(d.tf, 2 * P); DALLOC(d.tu, 2 * P); DALLOC(d.tb_key, 2 * P); DALLOC(d.tb_list, 2 * P);
    DALLOC(d.sm_loc, M);
    DALLOC(d.t_col_best, 3 * B); DALLOC(d.t_col_mask, 3 * B);
    const size_t items = std::max(2 * P, 3 * B);
    DALLOC(d.tk_a, items); DALLOC(d.tk_b, items);
    DALLOC(d.tv_a, std::max(2 * B, M)); DALLOC(d.tv_b, std::max(2 * B, M));
    DALLOC(d.tk64_a, 2 * B); DALLOC(d.tk64_b, 2 * B);
    d.t_sort_bytes = tile_sort_bytes(uint32_t(B), uint32_t(items));
    if (!d.t_sort_bytes) return fail(PXG_ERR_HIP, "tiled solve order: sizing the sorts failed");
    uint8_t* tmp = nullptr;
    DALLOC(tmp, d.t_sort_bytes);
    d.t_sort_tmp = tmp;
    return PXG_OK;
}

static uint32_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return uint32_t(p);
}

static int alloc_shapes(pxg_world* w, uint32_t n) {
    if (w->cap_s) {
        if (n > w->cap_s) return fail(PXG_ERR_CAPACITY, "pxg_set_shapes: more shapes than the world's capacity");
        return PXG_OK;
    }
    uint32_t cap = std::max(n, w->cfg.max_shapes);
    if (cap > kMaxShapes) return fail(PXG_ERR_CAPACITY, "pxg_set_shapes: at most " + std::to_string(kMaxShapes) + " shapes per world");
    w->cap_s = cap;
    w->cap_p = w->cfg.max_pairs ? w->cfg.max_pairs : 24u * cap + 4096u;
    w->cap_m = w->cfg.max_manifolds ? w->cfg.max_manifolds : std::min(w->cap_p, 12u * cap + 4096u);
    if (w->cap_p > (1u << 20)) w->cap_p = 1u << 20;  // scan limit
    if (w->cap_m > w->cap_p) w->cap_m = w->cap_p;
    DevState& d = w->d;
    d.cap_pairs = w->cap_p;
    d.cap_manifolds = w->cap_m;
    DALLOC(d.s_info, cap); DALLOC(d.s_lc, cap); DALLOC(d.s_lr, cap); DALLOC(d.s_dim, cap); DALLOC(d.s_mat, cap);
    DALLOC(d.s_fmin, cap); DALLOC(d.s_fmax, cap); DALLOC(d.s_tmin, cap); DALLOC(d.s_tmax, cap);
    DALLOC(d.cell_count, kMaxCells); DALLOC(d.cell_start, kMaxCells); DALLOC(d.cell_cursor, kMaxCells);
    DALLOC(d.ci_fmin, cap); DALLOC(d.ci_fmax, cap); DALLOC(d.ci_tmin, cap); DALLOC(d.ci_tmax, cap);
    DALLOC(d.large_list, kMaxLarge);
    DALLOC(d.pair_count, size_t(kPairLanes) * cap); DALLOC(d.pair_off, size_t(kPairLanes) * cap);
    size_t P = w->cap_p, M = w->cap_m;
    DALLOC(d.pairs, P); DALLOC(d.pm_valid, P); DALLOC(d.pm_off, P);
    DALLOC(d.pm_ab, P); DALLOC(d.pm_nf, P); DALLOC(d.pm_meta, P); DALLOC(d.pm_p, 4 * P); DALLOC(d.pm_q, 4 * P);
    DALLOC(d.ev_keys, P); DALLOC(d.ev_snap[0], P); DALLOC(d.ev_snap[1], P);
    d.ev_sort_bytes = events_sort_bytes(w->cap_p);
    if (!d.ev_sort_bytes) return fail(PXG_ERR_HIP, "pxg_set_shapes: sizing the contact-event sort failed");
    {
        uint8_t* tmp = nullptr;
        DALLOC(tmp, d.ev_sort_bytes);
        d.ev_sort_tmp = tmp;
    }
    DALLOC(d.ev_code, 2 * P); DALLOC(d.ev_boff, 12 * ((P + 1023) / 1024) + 12);
    d.cap_events = w->ev_cap_req ? w->ev_cap_req : uint32_t(std::min<size_t>(2 * M + 1024, 0xFFFFFFFFull));
    DALLOC(d.ev_out, d.cap_events);
    DALLOC(d.m_ab, M); DALLOC(d.m_shapes, M); DALLOC(d.m_nf, M); DALLOC(d.m_meta, M); DALLOC(d.m_p, 4 * M);
    DALLOC(d.m_q, 4 * M); DALLOC(d.m_colour, M);
    DALLOC(d.adj, 2 * M); DALLOC(d.m_pred, 2 * M);
    size_t nblk = (P + kBlock - 1) / kBlock;
    DALLOC(d.col_table, col_table_words(nblk)); DALLOC(d.col_table_off, size_t(kSortBuckets) * nblk + 16);
    DALLOC(d.order, M); DALLOC(d.ready_list, M); DALLOC(d.ready_cnt, 2);
    DALLOC(d.sm_ab, M); DALLOC(d.sm_nf, M); DALLOC(d.sm_cnt, M);
    DALLOC(d.sp_pos, 4 * M); DALLOC(d.sp_aux, 4 * M); DALLOC(d.sp_imp, 4 * M); DALLOC(d.sp_row, 4 * M);
    uint32_t cache_cap = next_pow2(8ull * M);
    d.cache_mask = cache_cap - 1;
    for (int t = 0; t < 2; ++t) {
        DALLOC(d.ck_keys[t], cache_cap); DALLOC(d.ck_ord[t], cache_cap); DALLOC(d.ck_val[t], cache_cap);
        launch_fill_u64(d.ck_keys[t], ~0ull, cache_cap, w->stream);
        launch_fill_u64(d.ck_ord[t], 0ull, cache_cap, w->stream);
    }
    d.cur = 0;
    return PXG_OK;
}

// ---------------------------------------------------------------------------------------------
// uploads and downloads
// ---------------------------------------------------------------------------------------------
// Between the two step phases (the on_substep seam, world.h:1299) only the state of the same
// bodies may change: the count and every body's type must stay, and the world inertia is kept.
extern "C" int pxg_set_bodies(pxg_world* w, const pxg_body* bodies, uint32_t n) {
    if (!w || (!bodies && n)) return fail(PXG_ERR_INVALID, "pxg_set_bodies: null argument");
    const bool seam = w->phase != 0;
    if (seam) {
        if (n != w->n_bodies) return fail(PXG_ERR_STATE, "pxg_set_bodies: the body count cannot change between the two step phases");
        for (uint32_t i = 0; i < n; ++i)
            if (bodies[i].type != w->body_types[i])
                return fail(PXG_ERR_STATE, "pxg_set_bodies: body types cannot change between the two step phases");
    }
    BodySoA soa;
    if (!pack_bodies(bodies, n, soa)) return fail(PXG_ERR_INVALID, "pxg_set_bodies: bad body type");
    if (!seam) {
        int r = alloc_bodies(w, n);
        if (r != PXG_OK) return r;
        w->body_types.resize(n);
        for (uint32_t i = 0; i < n; ++i) w->body_types[i] = bodies[i].type;
    }
    DevState& d = w->d;
    hipStream_t s = w->stream;
    if (n) {
        HIPCHK(hipMemcpyAsync(d.b_pos, soa.pos.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.b_rot, soa.rot.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.b_vel, soa.vel.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.b_angv, soa.angv.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.b_force, soa.force.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.b_torque, soa.torque.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.b_iil, soa.iil.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.b_prev_pos, soa.ppos.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.b_prev_rot, soa.prot.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.b_flags, soa.flags.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    }
    if (!seam) {
        d.nb = n;
        w->n_bodies = n;
        launch_init_bodies(d, s);
    }
    w->body_gen.resize(n);
    w->body_alive.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        w->body_gen[i] = bodies[i].generation;
        w->body_alive[i] = bodies[i].alive ? 1 : 0;
    }
    w->soft_resolve = true;
    HIPCHK(hipStreamSynchronize(s));  // host staging buffers go out of scope
    if (!seam && (!w->joints.empty() || !w->ignored.empty())) return rebuild_joints(w);  // body types order the joint layers
    return PXG_OK;
}

extern "C" int pxg_get_bodies(pxg_world* w, pxg_body* out, uint32_t n) {
    if (!w || (!out && n)) return fail(PXG_ERR_INVALID, "pxg_get_bodies: null argument");
    if (n > w->n_bodies) return fail(PXG_ERR_INVALID, "pxg_get_bodies: n exceeds the body count");
    if (!n) return PXG_OK;
    BodySoA soa;
    soa.resize(n);
    std::vector<float4> iiw(3 * size_t(n));
    DevState& d = w->d;
    hipStream_t s = w->stream;
    HIPCHK(hipMemcpyAsync(soa.pos.data(), d.b_pos, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(soa.rot.data(), d.b_rot, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(soa.vel.data(), d.b_vel, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(soa.angv.data(), d.b_angv, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(soa.force.data(), d.b_force, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(soa.torque.data(), d.b_torque, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(soa.iil.data(), d.b_iil, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(iiw.data(), d.b_iiw, 3 * n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(soa.ppos.data(), d.b_prev_pos, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(soa.prot.data(), d.b_prev_rot, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(soa.flags.data(), d.b_flags, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    unpack_bodies(soa, iiw, n, out);
    for (uint32_t i = 0; i < n && i < w->body_gen.size(); ++i) out[i].generation = w->body_gen[i];
    return PXG_OK;
}

extern "C" int pxg_set_bodies_subset(pxg_world* w, const uint32_t* indices, const pxg_body* records, uint32_t n) {
    if (!w || (n && (!indices || !records))) return fail(PXG_ERR_INVALID, "pxg_set_bodies_subset: null argument");
    if (!n) return PXG_OK;
    std::vector<uint8_t> seen(w->n_bodies, 0);
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t i = indices[k];
        if (i >= w->n_bodies) return fail(PXG_ERR_INVALID, "pxg_set_bodies_subset: index beyond the body count");
        if (seen[i]) return fail(PXG_ERR_INVALID, "pxg_set_bodies_subset: an index appears twice");
        seen[i] = 1;
        if (records[k].type != w->body_types[i])
            return fail(PXG_ERR_STATE, "pxg_set_bodies_subset: a body's type changes (upload the table with pxg_set_bodies)");
    }
    BodySoA soa;
    if (!pack_bodies(records, n, soa)) return fail(PXG_ERR_INVALID, "pxg_set_bodies_subset: bad body type");
    // one staging upload: the 9 float4 fields (SoA over the n records), then the flags and indices
    const size_t f4 = size_t(9) * n * sizeof(float4), words = size_t(2) * n * sizeof(uint32_t);
    if (f4 + words > w->body_stage_bytes) {
        HIPCHK(hipStreamSynchronize(w->stream));
        dfree(w, w->body_stage);
        w->body_stage = nullptr;
        w->body_stage_bytes = 0;
        DALLOC(w->body_stage, f4 + words);
        w->body_stage_bytes = f4 + words;
    }
    std::vector<uint8_t> host(f4 + words);
    const std::vector<float4>* fields[9] = {&soa.pos, &soa.rot, &soa.vel, &soa.angv, &soa.force, &soa.torque,
                                            &soa.iil, &soa.ppos, &soa.prot};
    for (int f = 0; f < 9; ++f) std::memcpy(host.data() + size_t(f) * n * sizeof(float4), fields[f]->data(), n * sizeof(float4));
    std::memcpy(host.data() + f4, soa.flags.data(), n * sizeof(uint32_t));
    std::memcpy(host.data() + f4 + n * sizeof(uint32_t), indices, n * sizeof(uint32_t));
    HIPCHK(hipMemcpyAsync(w->body_stage, host.data(), host.size(), hipMemcpyHostToDevice, w->stream));
    const float4* rec = reinterpret_cast<const float4*>(w->body_stage);
    const uint32_t* flags = reinterpret_cast<const uint32_t*>(w->body_stage + f4);
    const uint32_t* idx = flags + n;
    // as pxg_set_bodies: world inverse inertia recomputed outside the step, kept between its phases
    launch_scatter_bodies(w->d, idx, rec, flags, n, w->phase == 0 ? 1 : 0, w->stream);
    for (uint32_t k = 0; k < n; ++k) {
        w->body_gen[indices[k]] = records[k].generation;
        w->body_alive[indices[k]] = records[k].alive ? 1 : 0;
    }
    w->soft_resolve = true;
    HIPCHK(hipStreamSynchronize(w->stream));  // the host staging buffer goes out of scope
    return PXG_OK;
}

extern "C" int pxg_get_poses(pxg_world* w, pxg_pose* out, uint32_t n) {
    if (!w || (!out && n)) return fail(PXG_ERR_INVALID, "pxg_get_poses: null argument");
    if (n > w->n_bodies) return fail(PXG_ERR_INVALID, "pxg_get_poses: n exceeds the body count");
    if (!n) return PXG_OK;
    std::vector<float4> pos(n), rot(n), ppos(n), prot(n);
    std::vector<uint32_t> flags(n);
    const DevState& d = w->d;
    hipStream_t s = w->stream;
    HIPCHK(hipMemcpyAsync(pos.data(), d.b_pos, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(rot.data(), d.b_rot, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(ppos.data(), d.b_prev_pos, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(prot.data(), d.b_prev_rot, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(flags.data(), d.b_flags, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < n; ++i) {
        pxg_pose& o = out[i];
        o.position[0] = pos[i].x; o.position[1] = pos[i].y; o.position[2] = pos[i].z;
        o.type = flags[i] & BF_TYPE;
        o.orientation[0] = rot[i].x; o.orientation[1] = rot[i].y; o.orientation[2] = rot[i].z; o.orientation[3] = rot[i].w;
        o.prev_position[0] = ppos[i].x; o.prev_position[1] = ppos[i].y; o.prev_position[2] = ppos[i].z;
        o.interpolation = (flags[i] & BF_INTERP_NONE) ? PXG_INTERPOLATE_NONE : PXG_INTERPOLATE;
        o.prev_orientation[0] = prot[i].x; o.prev_orientation[1] = prot[i].y; o.prev_orientation[2] = prot[i].z;
        o.prev_orientation[3] = prot[i].w;
    }
    return PXG_OK;
}

extern "C" int pxg_set_hulls(pxg_world* w, const pxg_hull* hulls, uint32_t n) {
    if (!w || (n && !hulls)) return fail(PXG_ERR_INVALID, "pxg_set_hulls: null argument");
    if (w->phase != 0) return fail(PXG_ERR_STATE, "pxg_set_hulls: not allowed between the two step phases");
    if (n < w->hull_refs) return fail(PXG_ERR_STATE, "pxg_set_hulls: the shapes use hull " + std::to_string(w->hull_refs - 1));
    std::vector<uint4> meta(n), meta2(n);
    std::vector<float4> lb(2 * size_t(n)), cen(n), v, pl;
    std::vector<uint2> face, edge;
    std::vector<uint32_t> fv;
    for (uint32_t k = 0; k < n; ++k) {
        const pxg_hull& h = hulls[k];
        const std::string at = "pxg_set_hulls: hull " + std::to_string(k) + ": ";
        if (!h.vertices || !h.face_planes || !h.face_starts || !h.face_vertices || (h.n_edges && !h.edges))
            return fail(PXG_ERR_INVALID, at + "null array");
        if (h.n_vertices < 4 || h.n_vertices > 64) return fail(PXG_ERR_INVALID, at + "4 to 64 vertices");
        if (h.n_faces < 4) return fail(PXG_ERR_INVALID, at + "at least 4 faces");
        meta[k] = make_uint4(uint32_t(v.size()), h.n_vertices, uint32_t(pl.size()), h.n_faces);
        meta2[k] = make_uint4(uint32_t(edge.size()), h.n_edges, 0u, 0u);
        lb[2 * k] = make_float4(h.local_min[0], h.local_min[1], h.local_min[2], 0.0f);
        lb[2 * k + 1] = make_float4(h.local_max[0], h.local_max[1], h.local_max[2], 0.0f);
        cen[k] = make_float4(h.centroid[0], h.centroid[1], h.centroid[2], h.volume);
        for (uint32_t i = 0; i < h.n_vertices; ++i)
            v.push_back(make_float4(h.vertices[3 * i], h.vertices[3 * i + 1], h.vertices[3 * i + 2], 0.0f));
        for (uint32_t f = 0; f < h.n_faces; ++f) {
            const uint32_t s = h.face_starts[f], e = h.face_starts[f + 1];
            if (e < s + 3 || e - s > 8) return fail(PXG_ERR_INVALID, at + "a face has fewer than 3 or more than 8 vertices");
            pl.push_back(make_float4(h.face_planes[4 * f], h.face_planes[4 * f + 1], h.face_planes[4 * f + 2], h.face_planes[4 * f + 3]));
            face.push_back(make_uint2(uint32_t(fv.size()), e - s));
            for (uint32_t i = s; i < e; ++i) {
                if (h.face_vertices[i] >= h.n_vertices) return fail(PXG_ERR_INVALID, at + "face vertex index out of range");
                fv.push_back(h.face_vertices[i]);
            }
        }
        for (uint32_t i = 0; i < h.n_edges; ++i) {
            if (h.edges[2 * i] >= h.n_vertices || h.edges[2 * i + 1] >= h.n_vertices)
                return fail(PXG_ERR_INVALID, at + "edge vertex index out of range");
            edge.push_back(make_uint2(h.edges[2 * i], h.edges[2 * i + 1]));
        }
    }
    HIPCHK(hipStreamSynchronize(w->stream));
    DevState& d = w->d;
    void* old[9] = {d.h_meta, d.h_meta2, d.h_lb, d.h_cen, d.h_v, d.h_pl, d.h_face, d.h_e, d.h_fv};
    for (void* p : old) dfree(w, p);
    d.h_meta = d.h_meta2 = nullptr;
    d.h_lb = d.h_cen = d.h_v = d.h_pl = nullptr;
    d.h_face = d.h_e = nullptr;
    d.h_fv = nullptr;
    d.n_hulls = 0;
    if (!n) return PXG_OK;
    DALLOC(d.h_meta, n); DALLOC(d.h_meta2, n); DALLOC(d.h_lb, 2 * size_t(n)); DALLOC(d.h_cen, n);
    DALLOC(d.h_v, v.size()); DALLOC(d.h_pl, pl.size()); DALLOC(d.h_face, face.size()); DALLOC(d.h_fv, fv.size());
    DALLOC(d.h_e, std::max<size_t>(edge.size(), 1));
    hipStream_t st = w->stream;
    HIPCHK(hipMemcpyAsync(d.h_meta, meta.data(), n * sizeof(uint4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.h_meta2, meta2.data(), n * sizeof(uint4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.h_lb, lb.data(), lb.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.h_cen, cen.data(), cen.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.h_v, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.h_pl, pl.data(), pl.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.h_face, face.data(), face.size() * sizeof(uint2), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.h_fv, fv.data(), fv.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    if (!edge.empty()) HIPCHK(hipMemcpyAsync(d.h_e, edge.data(), edge.size() * sizeof(uint2), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    d.n_hulls = n;
    return PXG_OK;
}

static int check_mesh_desc(const pxg_mesh_desc* m, const std::string& at) {
    if (!m) return fail(PXG_ERR_INVALID, at + "null mesh");
    if ((m->n_positions && !m->positions) || (m->n_indices && !m->indices)) return fail(PXG_ERR_INVALID, at + "null array");
    return PXG_OK;
}

extern "C" int pxg_cook_mesh(const pxg_mesh_desc* m, pxg_mesh_info* info, float* vertices, uint32_t* indices,
                             float* normals, uint32_t* adjacency) {
    int r = check_mesh_desc(m, "pxg_cook_mesh: ");
    if (r != PXG_OK) return r;
    pxh::CookedMesh c;
    if (!pxh::cook_mesh(m->positions, m->n_positions, m->indices, m->n_indices, c))
        return fail(PXG_ERR_INVALID, "pxg_cook_mesh: an index is out of range");
    if (info) {
        info->n_vertices = uint32_t(c.vertices.size());
        info->n_triangles = c.n_tris();
        info->n_nodes = uint32_t(c.nodes.size());
        info->non_manifold_edges = c.non_manifold;
        const float lo[3] = {c.bounds.lo.x, c.bounds.lo.y, c.bounds.lo.z}, hi[3] = {c.bounds.hi.x, c.bounds.hi.y, c.bounds.hi.z};
        std::memcpy(info->local_min, lo, sizeof(lo));
        std::memcpy(info->local_max, hi, sizeof(hi));
    }
    for (size_t i = 0; vertices && i < c.vertices.size(); ++i) {
        vertices[3 * i] = c.vertices[i].x; vertices[3 * i + 1] = c.vertices[i].y; vertices[3 * i + 2] = c.vertices[i].z;
    }
    if (indices) std::copy(c.tris.begin(), c.tris.end(), indices);
    for (size_t t = 0; normals && t < c.normals.size(); ++t) {
        normals[3 * t] = c.normals[t].x; normals[3 * t + 1] = c.normals[t].y; normals[3 * t + 2] = c.normals[t].z;
    }
    if (adjacency) std::copy(c.neighbours.begin(), c.neighbours.end(), adjacency);
    return PXG_OK;
}

extern "C" int pxg_set_meshes(pxg_world* w, const pxg_mesh_desc* meshes, uint32_t n) {
    if (!w || (n && !meshes)) return fail(PXG_ERR_INVALID, "pxg_set_meshes: null argument");
    if (w->phase != 0) return fail(PXG_ERR_STATE, "pxg_set_meshes: not allowed between the two step phases");
    if (n < w->mesh_refs) return fail(PXG_ERR_STATE, "pxg_set_meshes: the shapes use mesh " + std::to_string(w->mesh_refs - 1));
    std::vector<uint4> meta(n), tri, adj;
    std::vector<float4> lb(2 * size_t(n)), v, nrm, node_lo, node_hi;
    std::vector<uint2> node_tri;
    std::vector<uint32_t> order;
    for (uint32_t k = 0; k < n; ++k) {
        const std::string at = "pxg_set_meshes: mesh " + std::to_string(k) + ": ";
        int r = check_mesh_desc(&meshes[k], at);
        if (r != PXG_OK) return r;
        pxh::CookedMesh c;
        if (!pxh::cook_mesh(meshes[k].positions, meshes[k].n_positions, meshes[k].indices, meshes[k].n_indices, c))
            return fail(PXG_ERR_INVALID, at + "an index is out of range");
        if (c.n_tris() == 0) return fail(PXG_ERR_INVALID, at + "no triangle left after welding");
        if (c.depth > pxh::kTreeDepthMax) return fail(PXG_ERR_INVALID, at + "triangle tree too deep");
        meta[k] = make_uint4(uint32_t(v.size()), uint32_t(tri.size()), uint32_t(node_lo.size()), c.n_tris());
        lb[2 * k] = make_float4(c.bounds.lo.x, c.bounds.lo.y, c.bounds.lo.z, 0.0f);
        lb[2 * k + 1] = make_float4(c.bounds.hi.x, c.bounds.hi.y, c.bounds.hi.z, 0.0f);
        for (const pxh::P3& p : c.vertices) v.push_back(make_float4(p.x, p.y, p.z, 0.0f));
        for (uint32_t t = 0; t < c.n_tris(); ++t) {
            tri.push_back(make_uint4(c.tris[3 * t], c.tris[3 * t + 1], c.tris[3 * t + 2], 0u));
            adj.push_back(make_uint4(c.neighbours[3 * t], c.neighbours[3 * t + 1], c.neighbours[3 * t + 2], 0u));
            nrm.push_back(make_float4(c.normals[t].x, c.normals[t].y, c.normals[t].z, 0.0f));
        }
        for (const pxh::MeshNodeH& nd : c.nodes) {
            float l, rgt;
            std::memcpy(&l, &nd.left, sizeof(l));
            std::memcpy(&rgt, &nd.right, sizeof(rgt));
            node_lo.push_back(make_float4(nd.box.lo.x, nd.box.lo.y, nd.box.lo.z, l));
            node_hi.push_back(make_float4(nd.box.hi.x, nd.box.hi.y, nd.box.hi.z, rgt));
            node_tri.push_back(make_uint2(nd.first, nd.count));
        }
        order.insert(order.end(), c.order.begin(), c.order.end());
    }
    HIPCHK(hipStreamSynchronize(w->stream));
    DevState& d = w->d;
    void* old[10] = {d.me_meta, d.me_lb, d.me_v, d.me_tri, d.me_n, d.me_adj, d.me_node_lo, d.me_node_hi, d.me_node_tri, d.me_order};
    for (void* p : old) dfree(w, p);
    d.me_meta = d.me_tri = d.me_adj = nullptr;
    d.me_lb = d.me_v = d.me_n = d.me_node_lo = d.me_node_hi = nullptr;
    d.me_node_tri = nullptr;
    d.me_order = nullptr;
    d.n_meshes = 0;
    if (!n) return PXG_OK;
    DALLOC(d.me_meta, n); DALLOC(d.me_lb, 2 * size_t(n)); DALLOC(d.me_v, v.size()); DALLOC(d.me_tri, tri.size());
    DALLOC(d.me_n, nrm.size()); DALLOC(d.me_adj, adj.size()); DALLOC(d.me_node_lo, node_lo.size());
    DALLOC(d.me_node_hi, node_hi.size()); DALLOC(d.me_node_tri, node_tri.size()); DALLOC(d.me_order, order.size());
    hipStream_t st = w->stream;
    HIPCHK(hipMemcpyAsync(d.me_meta, meta.data(), meta.size() * sizeof(uint4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.me_lb, lb.data(), lb.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.me_v, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.me_tri, tri.data(), tri.size() * sizeof(uint4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.me_n, nrm.data(), nrm.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.me_adj, adj.data(), adj.size() * sizeof(uint4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.me_node_lo, node_lo.data(), node_lo.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.me_node_hi, node_hi.data(), node_hi.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.me_node_tri, node_tri.data(), node_tri.size() * sizeof(uint2), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d.me_order, order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    d.n_meshes = n;
    return PXG_OK;
}

extern "C" int pxg_set_shapes(pxg_world* w, const pxg_shape* shapes, uint32_t n) {
    if (!w || (!shapes && n)) return fail(PXG_ERR_INVALID, "pxg_set_shapes: null argument");
    if (!w->cap_b) return fail(PXG_ERR_STATE, "pxg_set_shapes: upload bodies first");
    if (w->phase != 0) return fail(PXG_ERR_STATE, "pxg_set_shapes: not allowed between the two step phases");
    int r = alloc_shapes(w, n);
    if (r != PXG_OK) return r;
    ShapeSoA soa;
    uint32_t refs[2] = {0, 0};
    const char* err = pack_shapes(shapes, n, w->n_bodies, [&](size_t i) { return w->body_types[shapes[i].owner] == PXG_STATIC; }, soa,
                                  w->d.n_hulls, w->d.n_meshes, refs);
    if (err) return fail(PXG_ERR_INVALID, std::string("pxg_set_shapes: ") + err);
    w->hull_refs = refs[0];
    w->mesh_refs = refs[1];
    DevState& d = w->d;
    hipStream_t st = w->stream;
    if (n) {
        HIPCHK(hipMemcpyAsync(d.s_info, soa.info.data(), n * sizeof(uint4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.s_lc, soa.lc.data(), n * sizeof(float4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.s_lr, soa.lr.data(), n * sizeof(float4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.s_dim, soa.dim.data(), n * sizeof(float4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.s_mat, soa.mat.data(), n * sizeof(float4), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(d.layer_mask, w->layer_masks, sizeof(w->layer_masks), hipMemcpyHostToDevice, st));
    d.ns = n;
    w->n_shapes = n;
    launch_init_proxies(d, params(w, w->cfg.fixed_dt), st);
    HIPCHK(hipStreamSynchronize(st));
    w->shapes_ready = true;
    return PXG_OK;
}

// ---------------------------------------------------------------------------------------------
// joints and ignored pairs
// ---------------------------------------------------------------------------------------------
// Re-derives everything that depends on the joint table and the body table: live flags, the
// two layerings and the no-contact pairs; uploads them.
static int rebuild_joints(pxg_world* w) {
    DevState& d = w->d;
    hipStream_t st = w->stream;
    const uint32_t n = uint32_t(w->joints.size());
    std::vector<uint8_t> live(n, 0);
    std::vector<uint64_t> keys(w->ignored);
    std::vector<uint4> meta(n);
    for (uint32_t j = 0; j < n; ++j) {
        const pxg_joint& jt = w->joints[j];
        bool in_range = jt.body_a < w->n_bodies && jt.body_b < w->n_bodies;
        live[j] = (jt.valid && jt.enabled && in_range) ? 1 : 0;
        meta[j] = make_uint4(jt.type, in_range ? jt.body_a : 0u, in_range ? jt.body_b : 0u, live[j]);
        if (jt.valid && !jt.collide_connected && in_range) {  // joint_blocks_collision_, world.h:1686-1694
            uint32_t lo = std::min(jt.body_a, jt.body_b), hi = std::max(jt.body_a, jt.body_b);
            keys.push_back((uint64_t(lo) << 32) | hi);
        }
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<uint32_t> of, orv;
    uint32_t Lf = layer_joint_list(w->joints.data(), n, live, w->body_types, false, of, w->jl_fwd);
    uint32_t Lr = layer_joint_list(w->joints.data(), n, live, w->body_types, true, orv, w->jl_rev);
    uint32_t L = std::max(Lf, Lr);
    w->jl_fwd.resize(L + 1, w->jl_fwd.back());
    w->jl_rev.resize(L + 1, w->jl_rev.back());
    if (n > w->cap_j) {
        dfree(w, d.j_meta); dfree(w, d.j_def); dfree(w, d.j_acc); dfree(w, d.j_order); dfree(w, d.j_layer_off);
        d.j_meta = nullptr; d.j_def = nullptr; d.j_acc = nullptr; d.j_order = nullptr; d.j_layer_off = nullptr;
        uint32_t cap = std::max(n, 2 * w->cap_j);
        DALLOC(d.j_meta, cap); DALLOC(d.j_def, 6 * size_t(cap)); DALLOC(d.j_acc, 2 * size_t(cap));
        DALLOC(d.j_order, 2 * size_t(cap)); DALLOC(d.j_layer_off, 2 * (size_t(cap) + 1));
        w->cap_j = cap;
    }
    if (keys.size() > w->cap_nc) {
        dfree(w, d.nc_keys);
        d.nc_keys = nullptr;
        uint32_t cap = std::max<uint32_t>(uint32_t(keys.size()), 2 * w->cap_nc);
        DALLOC(d.nc_keys, cap);
        w->cap_nc = cap;
    }
    std::vector<uint32_t> ord(2 * size_t(n)), offs(2 * (size_t(L) + 1));
    std::copy(of.begin(), of.end(), ord.begin());
    std::copy(orv.begin(), orv.end(), ord.begin() + n);
    std::copy(w->jl_fwd.begin(), w->jl_fwd.end(), offs.begin());
    std::copy(w->jl_rev.begin(), w->jl_rev.end(), offs.begin() + L + 1);
    if (n) {
        HIPCHK(hipMemcpyAsync(d.j_meta, meta.data(), n * sizeof(uint4), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.j_order, ord.data(), ord.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.j_layer_off, offs.data(), offs.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    }
    if (!keys.empty())
        HIPCHK(hipMemcpyAsync(d.nc_keys, keys.data(), keys.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    d.nj = n;
    d.nj_layers = L;
    d.n_nc = uint32_t(keys.size());
    return PXG_OK;
}

extern "C" int pxg_set_joints(pxg_world* w, const pxg_joint* joints, uint32_t n) {
    if (!w || (!joints && n)) return fail(PXG_ERR_INVALID, "pxg_set_joints: null argument");
    for (uint32_t j = 0; j < n; ++j) {
        const pxg_joint& jt = joints[j];
        if (jt.type > PXG_CONE_TWIST) return fail(PXG_ERR_INVALID, "pxg_set_joints: bad joint type");
        if (jt.body_a >= w->n_bodies || jt.body_b >= w->n_bodies)
            return fail(PXG_ERR_INVALID, "pxg_set_joints: body index beyond the body table");
        if (jt.body_a == jt.body_b) return fail(PXG_ERR_INVALID, "pxg_set_joints: a joint needs two different bodies");
    }
    w->joints.assign(joints, joints + n);
    int r = rebuild_joints(w);
    if (r != PXG_OK) return r;
    if (!n) return PXG_OK;
    std::vector<float4> def(6 * size_t(n)), acc(2 * size_t(n));
    for (uint32_t j = 0; j < n; ++j) pack_joint(joints[j], &def[6 * size_t(j)], &acc[2 * size_t(j)]);
    DevState& d = w->d;
    HIPCHK(hipMemcpyAsync(d.j_def, def.data(), def.size() * sizeof(float4), hipMemcpyHostToDevice, w->stream));
    HIPCHK(hipMemcpyAsync(d.j_acc, acc.data(), acc.size() * sizeof(float4), hipMemcpyHostToDevice, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    return PXG_OK;
}

extern "C" int pxg_get_joints(pxg_world* w, pxg_joint* out, uint32_t n) {
    if (!w || (!out && n)) return fail(PXG_ERR_INVALID, "pxg_get_joints: null argument");
    if (n > w->joints.size()) return fail(PXG_ERR_INVALID, "pxg_get_joints: more joints requested than the table holds");
    if (!n) return PXG_OK;
    std::vector<float4> acc(2 * size_t(n));
    HIPCHK(hipMemcpyAsync(acc.data(), w->d.j_acc, acc.size() * sizeof(float4), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    for (uint32_t j = 0; j < n; ++j) {
        unpack_joint_acc(&acc[2 * size_t(j)], w->joints[j]);
        out[j] = w->joints[j];
    }
    return PXG_OK;
}

extern "C" int pxg_set_ignored_pairs(pxg_world* w, const uint32_t* ab, uint32_t n) {
    if (!w || (!ab && n)) return fail(PXG_ERR_INVALID, "pxg_set_ignored_pairs: null argument");
    w->ignored.clear();
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t a = ab[2 * i], b = ab[2 * i + 1];
        if (a == b) continue;  // set_pair_collision ignores a == b (world.h:528)
        w->ignored.push_back((uint64_t(std::min(a, b)) << 32) | std::max(a, b));
    }
    return rebuild_joints(w);
}

extern "C" int pxg_set_layer_masks(pxg_world* w, const uint32_t masks[32]) {
    if (!w || !masks) return fail(PXG_ERR_INVALID, "pxg_set_layer_masks: null argument");
    std::memcpy(w->layer_masks, masks, sizeof(w->layer_masks));
    if (w->d.layer_mask) {
        HIPCHK(hipMemcpyAsync(w->d.layer_mask, w->layer_masks, sizeof(w->layer_masks), hipMemcpyHostToDevice, w->stream));
        HIPCHK(hipStreamSynchronize(w->stream));
    }
    return PXG_OK;
}

// ---------------------------------------------------------------------------------------------
// stepping
// ---------------------------------------------------------------------------------------------
static int read_counters(pxg_world* w) {
    HIPCHK(hipMemcpyAsync(w->h_cnt, w->d.cnt, sizeof(Counters), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    uint32_t err = w->h_cnt->err;
    if (err) {
        uint32_t zero = 0;
        HIPCHK(hipMemcpy(&w->d.cnt->err, &zero, sizeof(zero), hipMemcpyHostToDevice));
        std::string what;
        if (err & ERR_PAIRS) what += " pair buffer full;";
        if (err & ERR_LARGE) what += " too many large shapes;";
        if (err & ERR_COLOURS) what += " more than 64 colours;";
        if (err & ERR_MANIFOLDS) what += " manifold buffer full;";
        if (err & ERR_CACHE) what += " warm-start table full;";
        if (err & ERR_MESH_VERTICES) what += " more than 32 mesh vertices inside one box or hull;";
        if (err & ERR_BARRIER_BIT) {
            what += " grid barrier timed out;";
            HIPCHK(hipMemset(w->d.bar, 0, kBarWords * sizeof(uint32_t)));
        }
        return fail(PXG_ERR_CAPACITY, "device reported:" + what);
    }
    return PXG_OK;
}

// gravity, force integration, broadphase, narrowphase (world.h:1278-1297)
static int phase_collide(pxg_world* w, float h) {
    DevState& d = w->d;
    hipStream_t st = w->stream;
    StepParams p = params(w, h);
    w->results_pending = false;  // a new step starts: the last step's counters are reset
    launch_integrate_forces(d, p, st);
    launch_broadphase_count(d, p, st);
    int r = read_counters(w);  // sync 1: pair count (also reports the last step's device errors)
    if (r != PXG_OK) return r;
    uint32_t P = w->h_cnt->pairs;
    if (P > w->cap_p) return fail(PXG_ERR_CAPACITY, "pair buffer full (" + std::to_string(P) + " pairs)");
    launch_broadphase_write(d, p, st);
    launch_narrowphase(d, p, P, st);
    if (w->ev_on) {  // world.h:1964-1978
        HIPCHK(launch_events(d, P, w->ev_prev_n, w->ev_cur, st));
        w->ev_prev_n = P;
        w->ev_cur ^= 1;
    }
    if (P == 0) {
        uint32_t zero = 0;
        HIPCHK(hipMemcpyAsync(&d.cnt->manifolds, &zero, sizeof(zero), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    w->last_pairs = P;
    return PXG_OK;
}

// Copies the last step's counters into the host-side statistics (the persistent path leaves
// them on the device until a caller asks); reports device errors.
static int finish_results(pxg_world* w) {
    if (!w->results_pending) return PXG_OK;
    w->results_pending = false;
    int r = read_counters(w);
    if (r != PXG_OK) return r;
    const Counters& hc = *w->h_cnt;
    std::memcpy(w->colour_off, hc.colour_off, sizeof(w->colour_off));
    w->last_manifolds = hc.manifolds;
    w->last_ordered = hc.ordered;
    w->last_active = hc.active;
    w->last_colours = hc.colours;
    pxg_stats& sst = w->stats;
    sst.manifolds = hc.manifolds;
    sst.points = hc.points;
    sst.colours = hc.colours;
    sst.active_manifolds = hc.active;
    sst.velocity_passes = w->cfg.residual_mode ? hc.passes : w->cfg.velocity_iterations;
    float residual = 0.0f;
    std::memcpy(&residual, &hc.residual_bits, 4);
    sst.residual = residual;
    sst.tiled_workgroups = w->cfg.solve_order == PXG_ORDER_TILED
                               ? (hc.tile_wg & 0xFFFFu) | (std::min((hc.tile_need + 1023u) >> 10, 0xFFFFu) << 16)
                               : 0u;
    return PXG_OK;
}

// A cooperative launch can be refused while another process holds the device's cooperative
// resources; the step then continues on the eager path (same results, one launch per colour).
static bool cooperative_refused(hipError_t e) {
    return e == hipErrorIllegalState || e == hipErrorNotSupported || e == hipErrorCooperativeLaunchTooLarge;
}
static void drop_persistent(pxg_world* w, hipError_t e) {
    (void)hipGetLastError();
    w->coop_blocks = 0;
    w->coop_note = std::string("persistent kernels unavailable (") + hipGetErrorString(e) + "), using per-colour launches";
}

// Eager colouring: rounds launched from the host until every active manifold has a colour; the
// colour table is read back because it sizes the per-colour launches.
static int colour_eager(pxg_world* w, uint32_t bound) {
    DevState& d = w->d;
    hipStream_t st = w->stream;
    const bool reference = w->cfg.solve_order == PXG_ORDER_REFERENCE;
    if (reference) {
        launch_reference_chains(d, bound, st);
        launch_reference_layers(d, bound, 24, st);
    } else {
        launch_colour_rounds(d, bound, 1, 24, st);
    }
    uint32_t round = 25;
    launch_build_order(d, bound, st);
    int r = read_counters(w);
    if (r != PXG_OK) return r;
    for (int extra = 0; w->h_cnt->remaining != 0 && extra < 64; ++extra) {
        if (reference) launch_reference_layers(d, bound, 16, st);
        else launch_colour_rounds(d, bound, round, 16, st);
        round += 16;
        launch_build_order(d, bound, st);
        r = read_counters(w);
        if (r != PXG_OK) return r;
    }
    if (w->h_cnt->remaining != 0) return fail(PXG_ERR_DEVICE, "colouring did not converge");
    return PXG_OK;
}

// Eager solve over the colour table in w->h_cnt: one launch per colour per pass. `setup_done`:
// the solver arrays (and learned predictions) were already written this step.
static int solve_eager(pxg_world* w, const StepParams& p, uint32_t bound, bool setup_done) {
    DevState& d = w->d;
    hipStream_t st = w->stream;
    const pxg_config& c = w->cfg;
    const Counters& hc = *w->h_cnt;
    uint32_t C = hc.colours, active = hc.active;
    std::memcpy(w->colour_off, hc.colour_off, sizeof(w->colour_off));
    const uint32_t* co = w->colour_off;
    if (!setup_done) {
        launch_solver_setup(d, p, bound, st);
        if (d.mlp && !d.record) launch_apply_predictions(d, 4 * active, st);
    }
    uint32_t passes = 0;
    const uint32_t L = d.nj_layers;
    auto joint_pass = [&](uint32_t it) {
        const bool rev = (it & 1u) != 0u;
        const std::vector<uint32_t>& off = rev ? w->jl_rev : w->jl_fwd;
        for (uint32_t l = 0; l < L; ++l) launch_joint_velocity(d, p, rev ? 1 : 0, off[l], off[l + 1], st);
    };
    for (uint32_t k = 0; k < C; ++k) launch_warm_apply(d, co[k], co[k + 1], st);
    for (uint32_t l = 0; l < L; ++l) launch_joint_warm(d, w->jl_fwd[l], w->jl_fwd[l + 1], st);
    if (c.residual_mode) {
        uint32_t maxit = c.residual_max_iterations ? c.residual_max_iterations : 256;
        const uint32_t every = c.residual_check_interval ? c.residual_check_interval : 1u;
        while (passes < maxit) {
            for (uint32_t k = 0; k < C; ++k) velocity_launch(w, p, co[k], co[k + 1], 0, passes);
            joint_pass(passes);
            ++passes;
            if (passes % every != 0u && passes != maxit) continue;
            launch_residual(d, p, active, 0, st);
            HIPCHK(hipMemcpyAsync(&w->h_cnt->residual_bits, &d.cnt->residual_bits, 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            float residual = 0.0f;
            std::memcpy(&residual, &w->h_cnt->residual_bits, 4);
            if (residual <= c.residual_tolerance) break;
        }
        launch_set_passes(d, passes, st);
    } else {
        for (uint32_t it = 0; it < c.velocity_iterations; ++it) {
            for (uint32_t k = 0; k < C; ++k) velocity_launch(w, p, co[k], co[k + 1], 0, it);
            joint_pass(it);
        }
    }
    for (uint32_t it = 0; it < c.relax_iterations; ++it) {
        for (uint32_t k = 0; k < C; ++k) velocity_launch(w, p, co[k], co[k + 1], 1, it);
        joint_pass(it);
    }
    launch_position_setup(d, st);
    for (uint32_t it = 0; it < c.position_iterations; ++it)
        for (uint32_t k = 0; k < C; ++k) launch_position(d, p, co[k], co[k + 1], st);
    for (uint32_t l = 0; l < L; ++l) launch_joint_position(d, p, w->jl_fwd[l], w->jl_fwd[l + 1], st);
    return PXG_OK;
}

// colouring, solve (dynamics/solver.h:914-958), islands/sleep, integration (world.h:1301-1309).
// The persistent path enqueues the whole phase without waiting on the device; the eager path
// (one launch per colour) reads the colour table back to size its launches.
static int phase_solve(pxg_world* w, float h) {
    DevState& d = w->d;
    hipStream_t st = w->stream;
    StepParams p = params(w, h);
    const pxg_config& c = w->cfg;
    uint32_t P = w->last_pairs;
    uint32_t bound = P ? P : 1;
    d.cur ^= 1;
    launch_wake_jointed(d, st);  // solve() starts with wake_jointed_islands (solver.h:916)
    launch_colour_init(d, bound, st);
    w->stats.pairs = P;
    w->stats.steps += 1;
    bool persistent = w->coop_blocks > 0 && !w->eager;
    // the tiled order: contact solves of worlds without joints (joint layers run between the
    // colours of one grid-wide order)
    bool tiled = persistent && c.solve_order == PXG_ORDER_TILED && d.nj == 0 && d.nb > 0;
    if (tiled) {
        int r = ensure_tiled(w);
        if (r != PXG_OK) return r;
    }
    if (persistent) {
        uint32_t nb = std::max(1u, std::min<uint32_t>(w->coop_blocks, (bound + kPersistBlock - 1) / kPersistBlock));
        nb = std::min(nb, w->colour_blocks);
        // Fresh colouring gives the manifolds with more points the low colours (see colour_key):
        // each later colour then holds only one-point manifolds and its phase is shorter, which
        // outweighs the extra colouring rounds when the solve runs hundreds of passes.
        const bool fresh = w->colour_fresh < 0 ? c.residual_mode != 0 : w->colour_fresh > 0;
        const uint32_t mode = c.solve_order == PXG_ORDER_REFERENCE ? 1u : (tiled ? 3u : (fresh ? 2u : 0u));
        hipError_t e = hipSuccess;
        DevState dc = d;
        if (tiled) {
            e = launch_tile_setup(d, bound, st);
            if (e != hipSuccess) return fail(PXG_ERR_HIP, std::string("tiled order: ") + hipGetErrorString(e));
            dc.col_best = d.t_col_best;  // one mask word per (phase, body)
            dc.col_mask = d.t_col_mask;
        }
        e = launch_colour_persistent(dc, 1024, nb, mode, st);
        if (e != hipSuccess) {
            if (!cooperative_refused(e)) return fail(PXG_ERR_HIP, std::string("colouring launch: ") + hipGetErrorString(e));
            drop_persistent(w, e);
            persistent = false;
            tiled = false;
        } else if (tiled) {
            e = launch_tile_order(d, bound, st);
            if (e != hipSuccess) return fail(PXG_ERR_HIP, std::string("tiled order: ") + hipGetErrorString(e));
        } else {
            launch_build_order(d, bound, st);
        }
    }
    if (!persistent) {
        int r = colour_eager(w, bound);
        if (r != PXG_OK) return r;
    }
    bool solved = false;
    if (persistent) {
        launch_solver_setup(d, p, bound, st);
        if (tiled) launch_tile_slots(d, bound, st);
        if (d.mlp && !d.record) launch_apply_predictions(d, 4 * bound, st);
        // the widest colour is at most the manifold count; threads loop when a colour is wider
        uint32_t snb = std::max(1u, std::min<uint32_t>(w->coop_blocks, (bound + kPersistBlock - 1) / kPersistBlock));
        snb = std::min(snb, w->solve_blocks);
        // one XCD: launch on all of them (blocks go to XCDs in turn); the blocks that land on one
        // carry on, the rest leave at once
        // the tiled solve: one workgroup per tile (on one XCD when every XCD can take them)
        if (tiled) snb = kTiles;
        const bool one_xcd = w->solve_one_xcd && static_cast<uint32_t>(w->coop_blocks) >= kXcds * snb;
        if (one_xcd) snb *= kXcds;
        // residual mode without joints: split the manifolds for the residual computed during the
        // tail colours
        const bool ready = c.residual_mode && d.nj == 0 && !tiled;
        if (ready) launch_ready_split(d, bound, st);
        uint32_t cfgv[8] = {c.velocity_iterations, c.relax_iterations, c.position_iterations, c.residual_mode,
                            c.residual_max_iterations ? c.residual_max_iterations : 256u, one_xcd ? 1u : 0u,
                            tiled ? (w->tiled_lds ? 1u : 0u) : (ready ? 1u : 0u),
                            c.residual_check_interval ? c.residual_check_interval : 1u};
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (w->profiling) {
            e0 = take_event(w);
            e1 = take_event(w);
            (void)hipEventRecord(e0, st);
        }
        hipError_t e = tiled ? launch_solve_tiled(d, p, cfgv, c.residual_tolerance, snb, st)
                             : launch_solve_persistent(d, p, cfgv, c.residual_tolerance, snb, st);
        if (e == hipSuccess) {
            solved = true;
            if (w->profiling) {
                (void)hipEventRecord(e1, st);
                w->ev_pending.push_back({e0, e1});
                launch_solve_bytes(d, cfgv, bound, st);
            }
        } else {
            if (!cooperative_refused(e)) return fail(PXG_ERR_HIP, std::string("solve launch: ") + hipGetErrorString(e));
            drop_persistent(w, e);
            int r = read_counters(w);  // the colour table for the per-colour launches
            if (r != PXG_OK) return r;
            r = solve_eager(w, p, bound, true);
            if (r != PXG_OK) return r;
            solved = true;
        }
    }
    if (!solved) {
        int r = solve_eager(w, p, bound, false);
        if (r != PXG_OK) return r;
    }
    launch_solver_finish(d, bound, st);
    launch_islands_sleep(d, p, bound, st);
    launch_integrate_velocities(d, p, st);
    launch_joints_post(d, p, c.position_iterations, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PXG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    w->results_pending = true;
    return PXG_OK;
}

extern "C" int pxg_step_phase(pxg_world* w, uint32_t phase, float h) {
    if (!w) return fail(PXG_ERR_INVALID, "pxg_step_phase: null handle");
    if (!w->shapes_ready) return fail(PXG_ERR_STATE, "pxg_step_phase: upload bodies and shapes first");
    if (!(h > 0.0f)) return fail(PXG_ERR_INVALID, "pxg_step_phase: h must be positive");
    if (phase == PXG_PHASE_FORCES_COLLIDE) {
        if (w->phase != 0) return fail(PXG_ERR_STATE, "pxg_step_phase: collision phase already run");
        int r = phase_collide(w, h);
        if (r != PXG_OK) return r;
        w->phase = 1;
        return PXG_OK;
    }
    if (phase == PXG_PHASE_SOLVE_INTEGRATE) {
        if (w->phase != 1) return fail(PXG_ERR_STATE, "pxg_step_phase: run the collision phase first");
        w->phase = 0;
        return phase_solve(w, h);
    }
    return fail(PXG_ERR_INVALID, "pxg_step_phase: unknown phase");
}

extern "C" int pxg_step_fixed(pxg_world* w, float h) {
    int r = pxg_step_phase(w, PXG_PHASE_FORCES_COLLIDE, h);
    if (r != PXG_OK) return r;
    return pxg_step_phase(w, PXG_PHASE_SOLVE_INTEGRATE, h);
}

// step(): accumulate the clamped frame time and run fixed substeps (world.h:1229-1240)
extern "C" int pxg_step(pxg_world* w, float dt, uint32_t* substeps_run) {
    if (!w) return fail(PXG_ERR_INVALID, "pxg_step: null handle");
    int rc = pxg_clear_events(w);  // world.h:1230
    if (rc != PXG_OK) return rc;
    float frame = std::min(dt, 0.25f);
    w->accumulator += frame;
    float h = w->cfg.fixed_dt;
    uint32_t n = 0;
    while (w->accumulator >= h && n < w->cfg.max_substeps) {
        int r = pxg_step_fixed(w, h);
        if (r != PXG_OK) return r;
        w->accumulator -= h;
        ++n;
    }
    if (substeps_run) *substeps_run = n;
    return step_soft(w, frame, n);
}

extern "C" int pxg_step_soft(pxg_world* w, float frame, uint32_t substeps_run) {
    if (!w) return fail(PXG_ERR_INVALID, "pxg_step_soft: null handle");
    if (w->phase != 0) return fail(PXG_ERR_STATE, "pxg_step_soft: not allowed between the two step phases");
    return step_soft(w, std::min(frame, 0.25f), substeps_run);
}

// ---------------------------------------------------------------------------------------------
// contact events (world.h:744-748, 1998-2018)
// ---------------------------------------------------------------------------------------------
extern "C" int pxg_clear_events(pxg_world* w) {
    if (!w) return fail(PXG_ERR_INVALID, "pxg_clear_events: null handle");
    if (!w->d.cnt) return PXG_OK;
    launch_fill_u32(&w->d.cnt->ev_n, 0u, 2, w->stream);  // ev_n, ev_dropped
    return PXG_OK;
}

extern "C" int pxg_set_events(pxg_world* w, uint32_t enabled, uint32_t capacity) {
    if (!w) return fail(PXG_ERR_INVALID, "pxg_set_events: null handle");
    bool on = enabled != 0;
    if (on && !w->ev_on && w->d.cnt) {  // no snapshot was kept while events were off
        w->ev_prev_n = 0;
        launch_fill_u32(w->d.cnt->ev_snap_n, 0u, 2, w->stream);
    }
    w->ev_on = on;
    if (capacity && capacity != w->d.cap_events && w->d.ev_out) {
        HIPCHK(hipStreamSynchronize(w->stream));
        dfree(w, w->d.ev_out);
        w->d.ev_out = nullptr;
        w->d.cap_events = 0;
        DALLOC(w->d.ev_out, capacity);
        w->d.cap_events = capacity;
        int r = pxg_clear_events(w);
        if (r != PXG_OK) return r;
    }
    if (capacity) w->ev_cap_req = capacity;
    return PXG_OK;
}

extern "C" int pxg_get_events(pxg_world* w, pxg_contact_event* out, uint32_t cap, uint32_t* n) {
    if (!w || !n || (!out && cap)) return fail(PXG_ERR_INVALID, "pxg_get_events: null argument");
    *n = 0;
    if (!w->d.cnt || !w->d.ev_out) return PXG_OK;
    uint32_t hv[2] = {0u, 0u};
    HIPCHK(hipMemcpyAsync(hv, &w->d.cnt->ev_n, sizeof(hv), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    *n = hv[0];
    uint32_t k = std::min(cap, hv[0]);
    if (k) {
        static_assert(sizeof(pxg_contact_event) == sizeof(uint4), "event record layout");
        HIPCHK(hipMemcpy(out, w->d.ev_out, size_t(k) * sizeof(uint4), hipMemcpyDeviceToHost));
    }
    if (hv[1])
        return fail(PXG_ERR_CAPACITY, "pxg_get_events: " + std::to_string(hv[1]) + " events were dropped: the queue holds " +
                                          std::to_string(w->d.cap_events) + " (clear it with pxg_clear_events, or raise "
                                          "the capacity with pxg_set_events)");
    return PXG_OK;
}

// ---------------------------------------------------------------------------------------------
// scene raycasts (world.h:766-847)
// ---------------------------------------------------------------------------------------------
static QueryFilterDev query_filter(const pxg_query_filter* f) {
    QueryFilterDev q;
    q.layer_mask = f ? f->layer_mask : ~0u;
    q.include_triggers = f ? (f->include_triggers != 0 ? 1u : 0u) : 1u;
    q.ignore_body = f ? f->ignore_body : PXG_NO_BODY;
    return q;
}
// query buffers for n rays (n hits), and the per-shape records filled from the current poses
static int query_setup(pxg_world* w, uint32_t n) {
    if (!w->shapes_ready) return fail(PXG_ERR_STATE, "scene query: upload bodies and shapes first");
    QueryScratch& q = w->q;
    const size_t S = w->cap_s;
    if (!q.geom) {
        DALLOC(q.geom, 3 * S); DALLOC(q.info, S); DALLOC(q.bounds6, 6);
        DALLOC(q.skey, S); DALLOC(q.skey2, S); DALLOC(q.sidx, S); DALLOC(q.sidx2, S);
        uint8_t* rec = nullptr;
        DALLOC(rec, S * query_shape_record_bytes());
        q.sorted = rec;
        DALLOC(q.tile_box, 2 * ((S + 255) / 256) + 2);
        DALLOC(q.chunk_box, 2 * ((S + 63) / 64) + 2);
        DALLOC(w->q_count, 1);
    }
    if (n > w->q_cap || !q.sort_tmp) {
        HIPCHK(hipStreamSynchronize(w->stream));
        void* old[8] = {w->q_rays, w->q_hits, q.rkey, q.rkey2, q.ridx, q.ridx2, q.sort_tmp, nullptr};
        for (void* p : old) dfree(w, p);
        w->q_rays = w->q_hits = nullptr;
        q.rkey = q.rkey2 = q.ridx = q.ridx2 = nullptr;
        q.sort_tmp = nullptr;
        w->q_cap = 0;
        const uint32_t cap = std::max<uint32_t>(std::max(n, w->q_cap), 1024u);
        DALLOC(w->q_rays, 2 * size_t(cap));
        DALLOC(w->q_hits, 3 * size_t(cap));
        DALLOC(q.rkey, cap); DALLOC(q.rkey2, cap); DALLOC(q.ridx, cap); DALLOC(q.ridx2, cap);
        q.sort_bytes = query_sort_bytes(std::max<uint32_t>(cap, w->cap_s));
        if (!q.sort_bytes) return fail(PXG_ERR_HIP, "scene query: sizing the sort failed");
        uint8_t* tmp = nullptr;
        DALLOC(tmp, q.sort_bytes);
        q.sort_tmp = tmp;
        w->q_cap = cap;
    }
    launch_fill_u32(q.bounds6, 0xFFFFFFFFu, 3, w->stream);
    launch_fill_u32(q.bounds6 + 3, 0u, 3, w->stream);
    launch_query_prep(w->d, q.geom, q.info, q.bounds6, w->stream);
    return PXG_OK;
}
static_assert(sizeof(pxg_ray) == 2 * sizeof(float4) && sizeof(pxg_raycast_hit) == 3 * sizeof(float4),
              "query record layouts");

static int raycast_batch(pxg_world* w, const pxg_ray* rays, uint32_t n, const pxg_query_filter* filter, uint32_t any,
                         const char* what) {
    int r = query_setup(w, n);
    if (r != PXG_OK) return r;
    HIPCHK(hipMemcpyAsync(w->q_rays, rays, size_t(n) * sizeof(pxg_ray), hipMemcpyHostToDevice, w->stream));
    HIPCHK(launch_raycast(w->d, w->q, w->q_rays, n, query_filter(filter), any, w->q_hits, w->stream));
    (void)what;
    return PXG_OK;
}

extern "C" int pxg_raycast(pxg_world* w, const pxg_ray* rays, uint32_t n, const pxg_query_filter* filter,
                           pxg_raycast_hit* hits) {
    if (!w || (n && (!rays || !hits))) return fail(PXG_ERR_INVALID, "pxg_raycast: null argument");
    if (!n) return PXG_OK;
    int r = raycast_batch(w, rays, n, filter, 0u, "pxg_raycast");
    if (r != PXG_OK) return r;
    HIPCHK(hipMemcpyAsync(hits, w->q_hits, size_t(n) * sizeof(pxg_raycast_hit), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    return PXG_OK;
}

extern "C" int pxg_raycast_device(pxg_world* w, const pxg_ray* d_rays, uint32_t n, const pxg_query_filter* filter,
                                  pxg_raycast_hit* d_hits) {
    if (!w || (n && (!d_rays || !d_hits))) return fail(PXG_ERR_INVALID, "pxg_raycast_device: null argument");
    if (!n) return PXG_OK;
    int r = query_setup(w, n);  // the ray keys and order need room for n rays
    if (r != PXG_OK) return r;
    HIPCHK(launch_raycast(w->d, w->q, reinterpret_cast<const float4*>(d_rays), n, query_filter(filter), 0u,
                          reinterpret_cast<float4*>(d_hits), w->stream));
    return PXG_OK;
}

extern "C" int pxg_raycast_any(pxg_world* w, const pxg_ray* rays, uint32_t n, const pxg_query_filter* filter,
                               uint32_t* any) {
    if (!w || (n && (!rays || !any))) return fail(PXG_ERR_INVALID, "pxg_raycast_any: null argument");
    if (!n) return PXG_OK;
    int r = raycast_batch(w, rays, n, filter, 1u, "pxg_raycast_any");
    if (r != PXG_OK) return r;
    std::vector<pxg_raycast_hit> hits(n);
    HIPCHK(hipMemcpyAsync(hits.data(), w->q_hits, size_t(n) * sizeof(pxg_raycast_hit), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    for (uint32_t i = 0; i < n; ++i) any[i] = hits[i].hit;
    return PXG_OK;
}

extern "C" int pxg_raycast_all(pxg_world* w, const pxg_ray* ray, const pxg_query_filter* filter, pxg_raycast_hit* hits,
                               uint32_t cap, uint32_t* n) {
    if (!w || !ray || !n || (cap && !hits)) return fail(PXG_ERR_INVALID, "pxg_raycast_all: null argument");
    *n = 0;
    // every shape can be hit once: the hit buffer holds one record per shape
    int r = query_setup(w, std::max<uint32_t>(w->n_shapes, 1u));
    if (r != PXG_OK) return r;
    HIPCHK(hipMemcpyAsync(w->q_rays, ray, sizeof(pxg_ray), hipMemcpyHostToDevice, w->stream));
    launch_fill_u32(w->q_count, 0u, 1, w->stream);
    launch_raycast_all(w->d, w->q.geom, w->q.info, w->q_rays, query_filter(filter), w->q_hits, w->q_cap, w->q_count,
                       w->stream);
    uint32_t count = 0;
    HIPCHK(hipMemcpyAsync(&count, w->q_count, sizeof(count), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    count = std::min(count, w->q_cap);
    std::vector<pxg_raycast_hit> all(count);
    if (count)
        HIPCHK(hipMemcpy(all.data(), w->q_hits, size_t(count) * sizeof(pxg_raycast_hit), hipMemcpyDeviceToHost));
    // nearest first (world.h:840-842); equal distances by shape slot
    std::sort(all.begin(), all.end(), [](const pxg_raycast_hit& a, const pxg_raycast_hit& b) {
        return a.distance < b.distance || (a.distance == b.distance && a.shape_index < b.shape_index);
    });
    *n = count;
    for (uint32_t i = 0; i < count && i < cap; ++i) hits[i] = all[i];
    return PXG_OK;
}

static OverlapDev overlap_dev(const pxg_overlap_query& query) {
    OverlapDev q;
    q.type = query.type;
    q.r = query.radius;
    q.c = make_float4(query.center[0], query.center[1], query.center[2], 0.0f);
    q.he = make_float4(query.half_extents[0], query.half_extents[1], query.half_extents[2], 0.0f);
    q.q = make_float4(query.orientation[0], query.orientation[1], query.orientation[2], query.orientation[3]);
    q.a = make_float4(query.a[0], query.a[1], query.a[2], 0.0f);
    q.b = make_float4(query.b[0], query.b[1], query.b[2], 0.0f);
    return q;
}

extern "C" int pxg_overlap_batch(pxg_world* w, const pxg_overlap_query* queries, uint32_t n, const pxg_query_filter* filter,
                                 uint32_t* bodies, uint32_t cap, uint32_t* counts) {
    if (!w || (n && (!queries || !counts)) || (n && cap && !bodies))
        return fail(PXG_ERR_INVALID, "pxg_overlap_batch: null argument");
    if (!n) return PXG_OK;
    std::vector<OverlapDev> q(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (queries[i].type > PXG_CAPSULE)
            return fail(PXG_ERR_INVALID, "pxg_overlap_batch: query " + std::to_string(i) + " is not a sphere, box or capsule");
        q[i] = overlap_dev(queries[i]);
    }
    int r = query_setup(w, 1u);
    if (r != PXG_OK) return r;
    const size_t out_n = std::max<size_t>(size_t(n) * cap, 1);
    if (n > w->q_ov_cap || out_n > w->q_ov_out_cap) {
        HIPCHK(hipStreamSynchronize(w->stream));
        if (n > w->q_ov_cap) {
            dfree(w, w->q_ov);
            dfree(w, w->q_ov_counts);
            w->q_ov = nullptr;
            w->q_ov_counts = nullptr;
            w->q_ov_cap = 0;
            DALLOC(w->q_ov, n);
            DALLOC(w->q_ov_counts, n);
            w->q_ov_cap = n;
        }
        if (out_n > w->q_ov_out_cap) {
            dfree(w, w->q_ov_out);
            w->q_ov_out = nullptr;
            w->q_ov_out_cap = 0;
            DALLOC(w->q_ov_out, out_n);
            w->q_ov_out_cap = out_n;
        }
    }
    HIPCHK(hipMemcpyAsync(w->q_ov, q.data(), size_t(n) * sizeof(OverlapDev), hipMemcpyHostToDevice, w->stream));
    launch_overlap_batch(w->d, w->q.geom, w->q.info, w->q.chunk_box, w->q_ov, n, query_filter(filter), w->q_ov_out, cap, w->q_ov_counts,
                         w->stream);
    HIPCHK(hipMemcpyAsync(counts, w->q_ov_counts, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToHost, w->stream));
    if (cap)
        HIPCHK(hipMemcpyAsync(bodies, w->q_ov_out, size_t(n) * cap * sizeof(uint32_t), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    return PXG_OK;
}

extern "C" int pxg_overlap_batch_noalloc(pxg_world* w, const pxg_overlap_query* d_queries, uint32_t n,
                                         const pxg_query_filter* filter, uint32_t* d_bodies, uint32_t cap,
                                         uint32_t* d_counts) {
    if (!w || (n && (!d_queries || !d_counts)) || (n && cap && !d_bodies))
        return fail(PXG_ERR_INVALID, "pxg_overlap_batch_noalloc: null argument");
    if (!w->shapes_ready) return fail(PXG_ERR_STATE, "pxg_overlap_batch_noalloc: upload bodies and shapes first");
    if (w->d.ns > kNoAllocShapes)
        return fail(PXG_ERR_CAPACITY, "pxg_overlap_batch_noalloc: the world has more than 65536 shape slots");
    if (!w->noalloc_scratch) return fail(PXG_ERR_HIP, "pxg_overlap_batch_noalloc: no device address for the scratch");
    if (!n) return PXG_OK;
    unsigned char* base = static_cast<unsigned char*>(w->noalloc_scratch);
    float4* geom = reinterpret_cast<float4*>(base + kScrGeom);
    uint4* info = reinterpret_cast<uint4*>(base + kScrInfo);
    uint32_t* bounds6 = reinterpret_cast<uint32_t*>(base + kScrBounds);
    float4* chunk_box = reinterpret_cast<float4*>(base + kScrChunks);
    // everything the kernels read from the scratch is written first in this sequence: the bounds
    // by fill kernels, the geometry and records of every slot by the geometry pass, every chunk's
    // box by k_chunk_bounds
    launch_fill_u32(bounds6, 0xFFFFFFFFu, 3, w->stream);
    launch_fill_u32(bounds6 + 3, 0u, 3, w->stream);
    launch_query_prep(w->d, geom, info, bounds6, w->stream);
    launch_overlap_batch_records(w->d, geom, info, chunk_box, d_queries, n, query_filter(filter), d_bodies, cap, d_counts,
                                 w->stream);
    return PXG_OK;
}

extern "C" int pxg_overlap(pxg_world* w, const pxg_overlap_query* query, const pxg_query_filter* filter, uint32_t* bodies,
                           uint32_t cap, uint32_t* n) {
    if (!w || !query || !n || (cap && !bodies)) return fail(PXG_ERR_INVALID, "pxg_overlap: null argument");
    *n = 0;
    if (query->type > PXG_CAPSULE) return fail(PXG_ERR_INVALID, "pxg_overlap: the query shape is a sphere, box or capsule");
    // one flag per shape: the ray-hit buffer (one record per shape at least) holds them
    int r = query_setup(w, std::max<uint32_t>(w->n_shapes, 1u));
    if (r != PXG_OK) return r;
    const OverlapDev q = overlap_dev(*query);
    uint32_t* flags = reinterpret_cast<uint32_t*>(w->q_hits);
    launch_overlap(w->d, w->q.geom, w->q.info, q, query_filter(filter), flags, w->stream);
    std::vector<uint32_t> h(w->n_shapes);
    if (!h.empty())
        HIPCHK(hipMemcpyAsync(h.data(), flags, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    uint32_t count = 0;
    for (uint32_t f : h) {
        if (!f) continue;
        if (count < cap) bodies[count] = f - 1u;
        ++count;
    }
    *n = count;
    return PXG_OK;
}

// a query shape (sphere, box or capsule; local centre, local rotation, capsule axis and dimensions)
// posed at position / orientation, as the query kernels take it
static CastDev posed_probe(const pxg_shape& sh, const float position[3], const float orientation[4]) {
    CastDev c{};
    c.shape.type = int(sh.type);
    c.shape.caxis = sh.capsule_axis;
    c.shape.lc.x = sh.local_center[0];
    c.shape.lc.y = sh.local_center[1];
    c.shape.lc.z = sh.local_center[2];
    c.shape.radius = sh.radius;
    c.shape.lr.x = sh.local_rotation[0];
    c.shape.lr.y = sh.local_rotation[1];
    c.shape.lr.z = sh.local_rotation[2];
    c.shape.lr.w = sh.local_rotation[3];
    c.shape.he.x = sh.half_extents[0];
    c.shape.he.y = sh.half_extents[1];
    c.shape.he.z = sh.half_extents[2];
    c.shape.cr = sh.capsule_radius;
    c.shape.chh = sh.capsule_half_height;
    c.pos = make_float4(position[0], position[1], position[2], 0.0f);
    c.rot = make_float4(orientation[0], orientation[1], orientation[2], orientation[3]);
    return c;
}

extern "C" int pxg_shape_cast(pxg_world* w, const pxg_cast_query* cast, const pxg_query_filter* filter,
                              pxg_raycast_hit* hit) {
    if (!w || !cast || !hit) return fail(PXG_ERR_INVALID, "pxg_shape_cast: null argument");
    std::memset(hit, 0, sizeof(*hit));
    if (cast->shape.type > PXG_CAPSULE)
        return fail(PXG_ERR_INVALID, "pxg_shape_cast: the cast shape is a sphere, a box or a capsule");
    // per-shape records in the hit buffer (3 float4 per shape), the hit itself after them
    int r = query_setup(w, std::max<uint32_t>(w->n_shapes, 1u) + 1u);
    if (r != PXG_OK) return r;
    if (!w->q_best) DALLOC(w->q_best, 1);
    CastDev c = posed_probe(cast->shape, cast->position, cast->orientation);
    c.dir = make_float4(cast->direction[0], cast->direction[1], cast->direction[2], 0.0f);
    c.maxd = cast->max_distance;
    launch_fill_u64(w->q_best, ~0ull, 1, w->stream);
    float4* out = w->q_hits + 3 * size_t(w->q_cap - 1);
    launch_shape_cast(w->d, w->q.info, c, query_filter(filter), w->q_hits, w->q_best, out, w->stream);
    HIPCHK(hipMemcpyAsync(hit, out, sizeof(pxg_raycast_hit), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    return PXG_OK;
}

extern "C" int pxg_shape_cast_batch(pxg_world* w, const pxg_cast_query* casts, uint32_t n, const pxg_query_filter* filter,
                                    pxg_raycast_hit* hits) {
    if (!w || (n && (!casts || !hits))) return fail(PXG_ERR_INVALID, "pxg_shape_cast_batch: null argument");
    if (!n) return PXG_OK;
    std::vector<CastDev> c(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (casts[i].shape.type > PXG_CAPSULE)
            return fail(PXG_ERR_INVALID, "pxg_shape_cast_batch: cast " + std::to_string(i) +
                                             " is not a sphere, a box or a capsule");
        c[i] = posed_probe(casts[i].shape, casts[i].position, casts[i].orientation);
        c[i].dir = make_float4(casts[i].direction[0], casts[i].direction[1], casts[i].direction[2], 0.0f);
        c[i].maxd = casts[i].max_distance;
    }
    int r = query_setup(w, n);  // hits: 3 float4 per cast in the hit buffer
    if (r != PXG_OK) return r;
    if (n > w->q_casts_cap) {
        HIPCHK(hipStreamSynchronize(w->stream));
        dfree(w, w->q_casts);
        w->q_casts = nullptr;
        w->q_casts_cap = 0;
        DALLOC(w->q_casts, n);
        w->q_casts_cap = n;
    }
    HIPCHK(hipMemcpyAsync(w->q_casts, c.data(), size_t(n) * sizeof(CastDev), hipMemcpyHostToDevice, w->stream));
    HIPCHK(launch_shape_cast_batch(w->d, w->q, w->q_casts, n, query_filter(filter), w->q_hits, w->stream));
    HIPCHK(hipMemcpyAsync(hits, w->q_hits, size_t(n) * sizeof(pxg_raycast_hit), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    return PXG_OK;
}

extern "C" int pxg_compute_penetration(pxg_world* w, const pxg_shape* shape, const float position[3],
                                       const float orientation[4], const pxg_query_filter* filter, pxg_penetration* out,
                                       uint32_t cap, uint32_t* n) {
    if (!w || !shape || !position || !orientation || !n || (cap && !out))
        return fail(PXG_ERR_INVALID, "pxg_compute_penetration: null argument");
    *n = 0;
    if (shape->type > PXG_CAPSULE)
        return fail(PXG_ERR_INVALID, "pxg_compute_penetration: the query shape is a sphere, a box or a capsule");
    // up to one record (3 float4) per shape in the hit buffer; the count in the nearest-key word
    int r = query_setup(w, std::max<uint32_t>(w->n_shapes, 1u) + 1u);
    if (r != PXG_OK) return r;
    if (!w->q_best) DALLOC(w->q_best, 1);
    const CastDev c = posed_probe(*shape, position, orientation);
    uint32_t* count = reinterpret_cast<uint32_t*>(w->q_best);
    launch_fill_u64(w->q_best, 0ull, 1, w->stream);
    launch_penetration(w->d, w->q.info, c, query_filter(filter), w->q_hits, count, w->stream);
    uint32_t k = 0;
    HIPCHK(hipMemcpyAsync(&k, count, sizeof(k), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    std::vector<float4> rec(3 * size_t(k));
    if (k) {
        HIPCHK(hipMemcpyAsync(rec.data(), w->q_hits, rec.size() * sizeof(float4), hipMemcpyDeviceToHost, w->stream));
        HIPCHK(hipStreamSynchronize(w->stream));
    }
    // appended in arrival order: report them in shape-slot order
    std::vector<uint32_t> order(k);
    for (uint32_t i = 0; i < k; ++i) order[i] = i;
    auto slot = [&](uint32_t i) {
        uint32_t v;
        std::memcpy(&v, &rec[3 * size_t(i) + 1].w, sizeof(v));
        return v;
    };
    std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return slot(x) < slot(y); });
    const uint32_t m = std::min(cap, k);
    for (uint32_t i = 0; i < m; ++i) {
        const float4 a = rec[3 * size_t(order[i])], b = rec[3 * size_t(order[i]) + 1];
        pxg_penetration& o = out[i];
        o.normal[0] = a.x;
        o.normal[1] = a.y;
        o.normal[2] = a.z;
        o.depth = a.w;
        o.point[0] = b.x;
        o.point[1] = b.y;
        o.point[2] = b.z;
        o.shape_index = slot(order[i]);
        std::memcpy(&o.body, &rec[3 * size_t(order[i]) + 2].x, sizeof(o.body));
        o.reserved = 0u;
    }
    *n = k;
    return PXG_OK;
}

extern "C" int pxg_compute_penetration_batch(pxg_world* w, const pxg_cast_query* probes, uint32_t n,
                                             const pxg_query_filter* filter, pxg_penetration* out, uint32_t cap,
                                             uint32_t* counts) {
    static_assert(sizeof(pxg_penetration) == 10 * sizeof(float), "pxg_penetration is 10 words");
    if (!w || (n && (!probes || !counts)) || (n && cap && !out))
        return fail(PXG_ERR_INVALID, "pxg_compute_penetration_batch: null argument");
    if (!n) return PXG_OK;
    std::vector<CastDev> c(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (probes[i].shape.type > PXG_CAPSULE)
            return fail(PXG_ERR_INVALID, "pxg_compute_penetration_batch: probe " + std::to_string(i) +
                                             " is not a sphere, a box or a capsule");
        c[i] = posed_probe(probes[i].shape, probes[i].position, probes[i].orientation);
    }
    int r = query_setup(w, 1u);
    if (r != PXG_OK) return r;
    const size_t out_words = std::max<size_t>(size_t(n) * cap * 10u, 1);
    if (n > w->q_casts_cap || n > w->q_ov_cap || out_words > w->q_pen_cap) {
        HIPCHK(hipStreamSynchronize(w->stream));
        if (n > w->q_casts_cap) {
            dfree(w, w->q_casts);
            w->q_casts = nullptr;
            w->q_casts_cap = 0;
            DALLOC(w->q_casts, n);
            w->q_casts_cap = n;
        }
        if (n > w->q_ov_cap) {
            dfree(w, w->q_ov);
            dfree(w, w->q_ov_counts);
            w->q_ov = nullptr;
            w->q_ov_counts = nullptr;
            w->q_ov_cap = 0;
            DALLOC(w->q_ov, n);
            DALLOC(w->q_ov_counts, n);
            w->q_ov_cap = n;
        }
        if (out_words > w->q_pen_cap) {
            dfree(w, w->q_pen);
            w->q_pen = nullptr;
            w->q_pen_cap = 0;
            DALLOC(w->q_pen, out_words);
            w->q_pen_cap = out_words;
        }
    }
    HIPCHK(hipMemcpyAsync(w->q_casts, c.data(), size_t(n) * sizeof(CastDev), hipMemcpyHostToDevice, w->stream));
    launch_penetration_batch(w->d, w->q.info, w->q.chunk_box, w->q_casts, n, query_filter(filter), w->q_pen, cap,
                             w->q_ov_counts, w->stream);
    HIPCHK(hipMemcpyAsync(counts, w->q_ov_counts, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToHost, w->stream));
    if (cap)
        HIPCHK(hipMemcpyAsync(out, w->q_pen, size_t(n) * cap * sizeof(pxg_penetration), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    return PXG_OK;
}

extern "C" int pxg_synchronize(pxg_world* w) {
    if (!w) return fail(PXG_ERR_INVALID, "pxg_synchronize: null handle");
    HIPCHK(hipStreamSynchronize(w->stream));
    return finish_results(w);
}

extern "C" int pxg_get_stats(pxg_world* w, pxg_stats* out) {
    if (!w || !out) return fail(PXG_ERR_INVALID, "pxg_get_stats: null argument");
    int r = finish_results(w);
    if (r != PXG_OK) return r;
    HIPCHK(hipMemcpyAsync(w->h_cnt, w->d.cnt, sizeof(Counters), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    w->stats.awake_bodies = w->h_cnt->awake;
    w->stats.learned_predictions = (w->d.mlp && !w->d.record) ? w->h_cnt->predicted : 0u;
    *out = w->stats;
    return PXG_OK;
}

extern "C" int pxg_get_pairs(pxg_world* w, uint32_t* ab, uint32_t cap, uint32_t* n) {
    if (!w || !n) return fail(PXG_ERR_INVALID, "pxg_get_pairs: null argument");
    *n = w->last_pairs;
    uint32_t k = std::min(cap, w->last_pairs);
    if (k && ab) {
        HIPCHK(hipMemcpyAsync(ab, w->d.pairs, k * sizeof(uint2), hipMemcpyDeviceToHost, w->stream));
        HIPCHK(hipStreamSynchronize(w->stream));
    }
    return PXG_OK;
}

extern "C" int pxg_get_solve_order(pxg_world* w, uint32_t* order, uint32_t cap, uint32_t* n, uint32_t* colour_offsets,
                                   uint32_t colour_cap, uint32_t* n_colours) {
    if (!w || !n || !n_colours) return fail(PXG_ERR_INVALID, "pxg_get_solve_order: null argument");
    {
        int fr = finish_results(w);
        if (fr != PXG_OK) return fr;
    }
    *n = w->last_ordered;
    *n_colours = w->last_colours;
    uint32_t k = std::min(cap, w->last_ordered);
    if (k && order) {
        HIPCHK(hipMemcpyAsync(order, w->d.order, k * sizeof(uint32_t), hipMemcpyDeviceToHost, w->stream));
        HIPCHK(hipStreamSynchronize(w->stream));
    }
    if (colour_offsets) {
        for (uint32_t c = 0; c <= w->last_colours && c < colour_cap; ++c) colour_offsets[c] = w->colour_off[c];
    }
    return PXG_OK;
}

extern "C" int pxg_get_manifolds(pxg_world* w, pxg_manifold* out, uint32_t cap, uint32_t* n) {
    if (!w || !n) return fail(PXG_ERR_INVALID, "pxg_get_manifolds: null argument");
    {
        int fr = finish_results(w);
        if (fr != PXG_OK) return fr;
    }
    uint32_t M = w->last_manifolds, O = w->last_ordered;
    *n = M;
    if (!out || !cap || !M) return PXG_OK;
    DevState& d = w->d;
    hipStream_t st = w->stream;
    std::vector<uint2> ab(M), sh(M);
    std::vector<float4> nf(M), p(4 * size_t(M)), q(4 * size_t(M)), aux(4 * size_t(O)), imp(4 * size_t(O));
    std::vector<uint4> meta(M);
    std::vector<uint32_t> order(O);
    HIPCHK(hipMemcpyAsync(ab.data(), d.m_ab, M * sizeof(uint2), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(sh.data(), d.m_shapes, M * sizeof(uint2), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(nf.data(), d.m_nf, M * sizeof(float4), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(meta.data(), d.m_meta, M * sizeof(uint4), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(p.data(), d.m_p, 4 * M * sizeof(float4), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(q.data(), d.m_q, 4 * M * sizeof(float4), hipMemcpyDeviceToHost, st));
    if (O) {
        HIPCHK(hipMemcpyAsync(order.data(), d.order, O * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(aux.data(), d.sp_aux, 4 * O * sizeof(float4), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(imp.data(), d.sp_imp, 4 * O * sizeof(float4), hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int64_t> slot(M, -1);
    for (uint32_t s = 0; s < O; ++s) slot[order[s]] = s;
    uint32_t k = std::min(cap, M);
    for (uint32_t m = 0; m < k; ++m) {
        pxg_manifold& o = out[m];
        std::memset(&o, 0, sizeof(o));
        o.a = ab[m].x; o.b = ab[m].y; o.shape_a = sh[m].x; o.shape_b = sh[m].y;
        o.normal[0] = nf[m].x; o.normal[1] = nf[m].y; o.normal[2] = nf[m].z; o.friction = nf[m].w;
        std::memcpy(&o.restitution, &meta[m].y, 4);
        o.count = meta[m].x & 255u;
        o.is_trigger = (meta[m].x >> 8) & 1u;
        o.valid = 1;
        for (int t = 0; t < 4; ++t) {
            pxg_contact& c = o.points[t];
            const float4& pp = p[4 * m + t];
            const float4& qq = q[4 * m + t];
            c.position[0] = pp.x; c.position[1] = pp.y; c.position[2] = pp.z; c.penetration = pp.w;
            c.separation = qq.x;
            std::memcpy(&c.feature_id, &qq.y, 4);
            if (slot[m] >= 0) {
                const float4& a = aux[4 * slot[m] + t];
                const float4& im = imp[4 * slot[m] + t];
                c.restitution_bias = a.y;
                c.normal_impulse = im.x; c.tangent_impulse[0] = im.y; c.tangent_impulse[1] = im.z;
            }
        }
    }
    return PXG_OK;
}

extern "C" int pxg_get_fat_aabbs(pxg_world* w, float* out6, uint32_t n) {
    if (!w || (!out6 && n)) return fail(PXG_ERR_INVALID, "pxg_get_fat_aabbs: null argument");
    if (n > w->n_shapes) return fail(PXG_ERR_INVALID, "pxg_get_fat_aabbs: n exceeds the shape count");
    if (!n) return PXG_OK;
    std::vector<float4> mn(n), mx(n);
    HIPCHK(hipMemcpyAsync(mn.data(), w->d.s_fmin, n * sizeof(float4), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipMemcpyAsync(mx.data(), w->d.s_fmax, n * sizeof(float4), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    for (uint32_t i = 0; i < n; ++i) {
        out6[6 * i] = mn[i].x; out6[6 * i + 1] = mn[i].y; out6[6 * i + 2] = mn[i].z;
        out6[6 * i + 3] = mx[i].x; out6[6 * i + 4] = mx[i].y; out6[6 * i + 5] = mx[i].z;
    }
    return PXG_OK;
}
extern "C" int pxg_set_fat_aabbs(pxg_world* w, const float* in6, uint32_t n) {
    if (!w || (!in6 && n)) return fail(PXG_ERR_INVALID, "pxg_set_fat_aabbs: null argument");
    if (n > w->n_shapes) return fail(PXG_ERR_INVALID, "pxg_set_fat_aabbs: n exceeds the shape count");
    if (!n) return PXG_OK;
    std::vector<float4> mn(n), mx(n);
    for (uint32_t i = 0; i < n; ++i) {
        mn[i] = make_float4(in6[6 * i], in6[6 * i + 1], in6[6 * i + 2], 0.0f);
        mx[i] = make_float4(in6[6 * i + 3], in6[6 * i + 4], in6[6 * i + 5], 0.0f);
    }
    HIPCHK(hipMemcpyAsync(w->d.s_fmin, mn.data(), n * sizeof(float4), hipMemcpyHostToDevice, w->stream));
    HIPCHK(hipMemcpyAsync(w->d.s_fmax, mx.data(), n * sizeof(float4), hipMemcpyHostToDevice, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    return PXG_OK;
}

extern "C" int pxg_get_warm_cache(pxg_world* w, uint64_t* keys, float* imp3, uint32_t cap, uint32_t* n) {
    if (!w || !n) return fail(PXG_ERR_INVALID, "pxg_get_warm_cache: null argument");
    *n = 0;
    if (!w->d.ck_keys[0]) return PXG_OK;
    size_t tc = size_t(w->d.cache_mask) + 1;
    std::vector<unsigned long long> k(tc);
    std::vector<float4> v(tc);
    int t = w->d.cur;
    HIPCHK(hipMemcpyAsync(k.data(), w->d.ck_keys[t], tc * 8, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipMemcpyAsync(v.data(), w->d.ck_val[t], tc * sizeof(float4), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    std::vector<std::pair<uint64_t, float4>> e;
    for (size_t i = 0; i < tc; ++i)
        if (k[i] != ~0ull) e.push_back({k[i], v[i]});
    std::sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    *n = uint32_t(e.size());
    for (uint32_t i = 0; i < cap && i < e.size(); ++i) {
        if (keys) keys[i] = e[i].first;
        if (imp3) { imp3[3 * i] = e[i].second.x; imp3[3 * i + 1] = e[i].second.y; imp3[3 * i + 2] = e[i].second.z; }
    }
    return PXG_OK;
}

static uint32_t host_hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}
static uint32_t host_hash64(uint64_t k) { return host_hash32(uint32_t(k) ^ host_hash32(uint32_t(k >> 32) + 0x9e3779b9u)); }

extern "C" int pxg_set_warm_cache(pxg_world* w, const uint64_t* keys, const float* imp3, uint32_t n) {
    if (!w || ((!keys || !imp3) && n)) return fail(PXG_ERR_INVALID, "pxg_set_warm_cache: null argument");
    if (!w->d.ck_keys[0]) return fail(PXG_ERR_STATE, "pxg_set_warm_cache: upload shapes first");
    size_t tc = size_t(w->d.cache_mask) + 1;
    if (n > tc / 2) return fail(PXG_ERR_CAPACITY, "pxg_set_warm_cache: too many entries");
    std::vector<unsigned long long> k(tc, ~0ull);
    std::vector<unsigned long long> o(tc, 0ull);
    std::vector<float4> v(tc, make_float4(0, 0, 0, 0));
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t idx = host_hash64(keys[i]) & w->d.cache_mask;
        while (k[idx] != ~0ull && k[idx] != keys[i]) idx = (idx + 1) & w->d.cache_mask;
        k[idx] = keys[i];
        o[idx] = 1ull;
        // w carries the manifold's colour of the step that wrote the entry: none for uploads
        const uint32_t no_colour = 0xFFFFFFFFu;
        float nc;
        std::memcpy(&nc, &no_colour, 4);
        v[idx] = make_float4(imp3[3 * i], imp3[3 * i + 1], imp3[3 * i + 2], nc);
    }
    int t = w->d.cur;
    HIPCHK(hipMemcpyAsync(w->d.ck_keys[t], k.data(), tc * 8, hipMemcpyHostToDevice, w->stream));
    HIPCHK(hipMemcpyAsync(w->d.ck_ord[t], o.data(), tc * 8, hipMemcpyHostToDevice, w->stream));
    HIPCHK(hipMemcpyAsync(w->d.ck_val[t], v.data(), tc * sizeof(float4), hipMemcpyHostToDevice, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    return PXG_OK;
}

// FNV-1a over alive bodies' position, orientation and velocities (world.h:1334-1359)
// ---------------------------------------------------------------------------------------------
// cloth and hair (world.h:552-700 tables, 1749-1957 stepping)
// ---------------------------------------------------------------------------------------------
namespace {

// util::compute_bounds over particle positions (cloth.h:320-327): glm::min / max in order
void particle_bounds(const pxg_particle* ps, uint32_t n, float mn[3], float mx[3]) {
    for (int k = 0; k < 3; ++k) {
        mn[k] = 3.402823466e+38f;
        mx[k] = -3.402823466e+38f;
    }
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            const float v = ps[i].position[k];
            if (v < mn[k]) mn[k] = v;
            if (mx[k] < v) mx[k] = v;
        }
}
// ClothSolver's bucket count (cloth_solver.h:96-100)
uint32_t cloth_buckets(size_t n) {
    uint32_t b = 16;
    while (b < n * 2u && b < (1u << 22)) b <<= 1;
    return b;
}
// [begin, end) pairs within n_con, each batch's constraints touching disjoint particles below np
const char* check_batches(const uint32_t* b, uint32_t nb, const pxg_distance_constraint* con, uint32_t n_con, uint32_t np,
                          std::vector<uint32_t>& stamp, uint32_t& tick) {
    for (uint32_t k = 0; k < nb; ++k) {
        const uint32_t begin = b[2 * k], end = b[2 * k + 1];
        if (begin > end || end > n_con) return "a batch is not a range of its constraints";
        ++tick;
        for (uint32_t i = begin; i < end; ++i) {
            const uint32_t a = con[i].a, c = con[i].b;
            if (a >= np || c >= np) return "a constraint's particle is out of range";
            if (stamp[a] == tick || (c != a && stamp[c] == tick)) return "a batch's constraints share a particle";
            stamp[a] = stamp[c] = tick;
        }
    }
    return nullptr;
}
inline float4 f4(const float* v, float w) { return make_float4(v[0], v[1], v[2], w); }
inline float4 f4q(const float* q) { return make_float4(q[0], q[1], q[2], q[3]); }
inline float u2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline uint32_t f2u(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
void put_particle(const float4& p, const float4& q, const float4& v, pxg_particle& o) {
    o.position[0] = p.x; o.position[1] = p.y; o.position[2] = p.z; o.inv_mass = p.w;
    o.prev_position[0] = q.x; o.prev_position[1] = q.y; o.prev_position[2] = q.z;
    o.velocity[0] = v.x; o.velocity[1] = v.y; o.velocity[2] = v.z;
}

}  // namespace

template <class T>
static int soft_alloc(pxg_world* w, std::vector<void*>& bufs, T*& dst, const std::vector<T>& src) {
    const size_t n = std::max<size_t>(src.size(), 1);
    T* p = nullptr;
    int r = dalloc(w, p, n);
    if (r != PXG_OK) return r;
    bufs.push_back(p);
    dst = p;
    if (!src.empty()) HIPCHK(hipMemcpyAsync(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, w->stream));
    return PXG_OK;
}
template <class T>
static int soft_upload(pxg_world* w, std::vector<void*>& bufs, const T*& dst, const std::vector<T>& src) {
    T* p = nullptr;
    int r = soft_alloc(w, bufs, p, src);
    dst = p;
    return r;
}
#define SOFT_ALLOC(bufs, dst, src)                      \
    do {                                                \
        int r_ = soft_alloc(w, (bufs), (dst), (src));   \
        if (r_ != PXG_OK) return r_;                    \
    } while (0)
#define SOFT_UPLOAD(bufs, dst, src)                     \
    do {                                                \
        int r_ = soft_upload(w, (bufs), (dst), (src));  \
        if (r_ != PXG_OK) return r_;                    \
    } while (0)

static int soft_lds(pxg_world* w, uint32_t np) {
    const size_t bytes = size_t(np) * sizeof(float4);
    if (bytes <= w->soft_lds) return PXG_OK;
    HIPCHK(soft_lds_limit(bytes));
    w->soft_lds = bytes;
    return PXG_OK;
}

// anchors' and collide bodies' (index, generation) against the bodies as last uploaded
// (BodyId validity, world.h:1764-1772 and 1893-1899)
static int soft_resolve(pxg_world* w) {
    auto live = [w](uint32_t index, uint32_t gen) {
        return index < w->body_gen.size() && w->body_alive[index] && w->body_gen[index] == gen;
    };
    if (!w->cl_anchor_res.empty()) {
        for (size_t a = 0; a < w->cl_anchor_res.size(); ++a) {
            const uint32_t* id = &w->cl_anchor_ids[3 * a];
            w->cl_anchor_res[a] = id[0] && live(id[1], id[2]) ? id[1] : kNoBody;
        }
        HIPCHK(hipMemcpyAsync(const_cast<uint32_t*>(w->sd.cl_anchor_body), w->cl_anchor_res.data(),
                              w->cl_anchor_res.size() * sizeof(uint32_t), hipMemcpyHostToDevice, w->stream));
    }
    if (!w->st_collide_res.empty()) {
        for (size_t k = 0; k < w->st_collide_res.size(); ++k) {
            const uint32_t* id = &w->st_collide_ids[2 * k];
            w->st_collide_res[k] = live(id[0], id[1]) ? id[0] : kNoBody;
        }
        HIPCHK(hipMemcpyAsync(const_cast<uint32_t*>(w->sd.st_collide), w->st_collide_res.data(),
                              w->st_collide_res.size() * sizeof(uint32_t), hipMemcpyHostToDevice, w->stream));
    }
    HIPCHK(hipStreamSynchronize(w->stream));  // the host arrays may change on the next call
    w->soft_resolve = false;
    return PXG_OK;
}

extern "C" int pxg_set_cloths(pxg_world* w, const pxg_cloth* cloths, uint32_t n) {
    if (!w || (n && !cloths)) return fail(PXG_ERR_INVALID, "pxg_set_cloths: null argument");
    if (w->phase != 0) return fail(PXG_ERR_STATE, "pxg_set_cloths: not allowed between the two step phases");
    std::vector<ClothRec> rec(n);
    std::vector<SoftState> state(n);
    std::vector<float4> pos, prev, vel, anchor;
    std::vector<uint4> con, tether, tri;
    std::vector<uint2> batch;
    std::vector<uint32_t> ptri_off, ptri, ids;
    size_t n_bins = 0, n_tri = 0;
    std::vector<uint32_t> stamp;
    uint32_t tick = 0, max_np = 0;
    for (uint32_t ci = 0; ci < n; ++ci) {
        const pxg_cloth& c = cloths[ci];
        ClothRec& r = rec[ci];
        std::memset(&r, 0, sizeof(r));
        SoftState& st = state[ci];
        std::memset(&st, 0, sizeof(st));
        const std::string at = "pxg_set_cloths: cloth " + std::to_string(ci) + ": ";
        r.alive = c.alive && c.n_particles ? 1u : 0u;
        if (!r.alive) continue;
        const uint32_t np = c.n_particles;
        if (!c.particles) return fail(PXG_ERR_INVALID, at + "null particles");
        if (np > kSoftMaxParticles) return fail(PXG_ERR_CAPACITY, at + "more than 8192 particles");
        if ((c.n_stretch && !c.stretch) || (c.n_bend && !c.bend) || (c.n_stretch_batches && !c.stretch_batches) ||
            (c.n_bend_batches && !c.bend_batches) || (c.n_tethers && !c.tethers) || (c.n_anchors && !c.anchors) ||
            (c.n_triangle_indices && !c.triangles))
            return fail(PXG_ERR_INVALID, at + "null array");
        stamp.assign(np, 0u);
        tick = 0;
        if (const char* e = check_batches(c.stretch_batches, c.n_stretch_batches, c.stretch, c.n_stretch, np, stamp, tick))
            return fail(PXG_ERR_INVALID, at + e);
        if (const char* e = check_batches(c.bend_batches, c.n_bend_batches, c.bend, c.n_bend, np, stamp, tick))
            return fail(PXG_ERR_INVALID, at + e);
        max_np = std::max(max_np, np);
        r.p0 = uint32_t(pos.size());
        r.np = np;
        r.c0 = uint32_t(con.size());
        r.n_stretch = c.n_stretch;
        r.n_bend = c.n_bend;
        r.b0 = uint32_t(batch.size());
        r.n_sb = c.n_stretch_batches;
        r.n_bb = c.n_bend_batches;
        r.t0 = uint32_t(tether.size());
        r.n_tether = c.n_tethers;
        r.a0 = uint32_t(anchor.size() / 3);
        r.n_anchor = c.n_anchors;
        r.tr0 = uint32_t(n_tri);
        r.n_tri = c.n_triangle_indices / 3;
        r.pt0 = uint32_t(ptri_off.size());
        r.columns = c.columns;
        r.enabled = c.enabled ? 1u : 0u;
        const pxg_cloth_params& q = c.params;
        r.self_collision = q.self_collision ? 1u : 0u;
        r.substeps = q.substeps;
        r.iterations = q.iterations;
        r.layer = q.layer;
        r.stretch_compliance = q.stretch_compliance; r.bend_compliance = q.bend_compliance; r.damping = q.damping;
        r.thickness = q.thickness; r.friction = q.friction; r.gravity_scale = q.gravity_scale;
        r.max_velocity = q.max_velocity; r.wind_turbulence = q.wind_turbulence; r.air_drag = q.air_drag;
        r.air_lift = q.air_lift; r.self_distance = q.self_distance; r.sleep_threshold = q.sleep_threshold;
        r.sleep_time = q.sleep_time;
        for (int k = 0; k < 3; ++k) {
            r.ext[k] = q.external_acceleration[k];
            r.wind[k] = q.wind[k];
        }
        if (r.self_collision && r.self_distance > 1e-6f) {
            r.buckets = cloth_buckets(np);
            r.bins0 = uint32_t(n_bins);
            n_bins += 2 * size_t(np) + r.buckets + 1;
        }
        for (uint32_t i = 0; i < np; ++i) {
            const pxg_particle& p = c.particles[i];
            pos.push_back(f4(p.position, p.inv_mass));
            prev.push_back(f4(p.prev_position, p.inv_mass));
            vel.push_back(f4(p.velocity, 0.0f));
        }
        for (uint32_t i = 0; i < c.n_stretch; ++i)
            con.push_back(make_uint4(c.stretch[i].a, c.stretch[i].b, f2u(c.stretch[i].rest_length), f2u(c.stretch[i].lambda)));
        for (uint32_t i = 0; i < c.n_bend; ++i)
            con.push_back(make_uint4(c.bend[i].a, c.bend[i].b, f2u(c.bend[i].rest_length), f2u(c.bend[i].lambda)));
        for (uint32_t k = 0; k < c.n_stretch_batches; ++k)
            batch.push_back(make_uint2(c.stretch_batches[2 * k], c.stretch_batches[2 * k + 1]));
        for (uint32_t k = 0; k < c.n_bend_batches; ++k)  // bend constraints follow the stretch ones
            batch.push_back(make_uint2(c.n_stretch + c.bend_batches[2 * k], c.n_stretch + c.bend_batches[2 * k + 1]));
        // tethers run in parallel unless one's anchor is moved by another tether, or two move one particle
        ++tick;
        bool serial = false;
        for (uint32_t t = 0; t < c.n_tethers; ++t) {
            const pxg_cloth_tether& te = c.tethers[t];
            if (te.particle >= np || te.anchor >= np) return fail(PXG_ERR_INVALID, at + "a tether's particle is out of range");
            if (stamp[te.particle] == tick) serial = true;
            stamp[te.particle] = tick;
            tether.push_back(make_uint4(te.particle, te.anchor, f2u(te.max_length), 0u));
        }
        for (uint32_t t = 0; t < c.n_tethers && !serial; ++t)
            if (stamp[c.tethers[t].anchor] == tick) serial = true;
        r.tether_serial = serial ? 1u : 0u;
        for (uint32_t a = 0; a < c.n_anchors; ++a) {
            const pxg_cloth_anchor& x = c.anchors[a];
            anchor.push_back(f4(x.world_position, u2f(x.particle)));
            anchor.push_back(f4(x.prev_world_position, 0.0f));
            anchor.push_back(f4(x.local_position, 0.0f));
            ids.push_back(x.has_body ? 1u : 0u);
            ids.push_back(x.body);
            ids.push_back(x.generation);
        }
        // per particle, its triangles in ascending order, once per corner it is
        std::vector<uint32_t> count(np + 1, 0u);
        for (uint32_t t = 0; t < r.n_tri; ++t)
            for (int k = 0; k < 3; ++k) {
                const uint32_t i = c.triangles[3 * t + k];
                if (i >= np) return fail(PXG_ERR_INVALID, at + "a triangle's particle is out of range");
                ++count[i];
            }
        const uint32_t base = uint32_t(ptri.size());
        std::vector<uint32_t> fill(np, 0u);
        uint32_t run = base;
        for (uint32_t i = 0; i < np; ++i) {
            ptri_off.push_back(run);
            fill[i] = run;
            run += count[i];
        }
        ptri_off.push_back(run);
        ptri.resize(run);
        for (uint32_t t = 0; t < r.n_tri; ++t) {
            const uint32_t a = c.triangles[3 * t], b = c.triangles[3 * t + 1], cc = c.triangles[3 * t + 2];
            tri.push_back(make_uint4(a, b, cc, 0u));
            ptri[fill[a]++] = t;
            ptri[fill[b]++] = t;
            ptri[fill[cc]++] = t;
        }
        n_tri += r.n_tri;
        st.awake = c.awake ? 1u : 0u;
        st.sleep_timer = c.sleep_timer;
        st.time = c.time;
        particle_bounds(c.particles, np, st.bmin, st.bmax);  // add_cloth recomputes them (world.h:555)
    }
    int rr = soft_lds(w, max_np);
    if (rr != PXG_OK) return rr;
    HIPCHK(hipStreamSynchronize(w->stream));
    for (void* p : w->cloth_bufs) dfree(w, p);
    w->cloth_bufs.clear();
    SoftDev& s = w->sd;
    s.n_cloths = 0;
    w->cl_rec.clear();
    w->cl_anchor_ids.clear();
    w->cl_anchor_res.clear();
    w->cl_max_np = 0;
    if (!n) return PXG_OK;
    std::vector<float4> share(std::max<size_t>(n_tri, 1));
    std::vector<uint32_t> bins(n_bins), cols(size_t(n) * kSoftMaxColliders), res(ids.size() / 3, kNoBody);
    SoftState* dstate = nullptr;
    SOFT_UPLOAD(w->cloth_bufs, s.cl, rec);
    SOFT_ALLOC(w->cloth_bufs, dstate, state);
    s.cl_state = dstate;
    SOFT_ALLOC(w->cloth_bufs, s.cl_pos, pos);
    SOFT_ALLOC(w->cloth_bufs, s.cl_prev, prev);
    SOFT_ALLOC(w->cloth_bufs, s.cl_vel, vel);
    SOFT_ALLOC(w->cloth_bufs, s.cl_con, con);
    SOFT_UPLOAD(w->cloth_bufs, s.cl_batch, batch);
    SOFT_UPLOAD(w->cloth_bufs, s.cl_tether, tether);
    SOFT_ALLOC(w->cloth_bufs, s.cl_anchor, anchor);
    SOFT_UPLOAD(w->cloth_bufs, s.cl_anchor_body, res);
    SOFT_UPLOAD(w->cloth_bufs, s.cl_tri, tri);
    SOFT_UPLOAD(w->cloth_bufs, s.cl_ptri_off, ptri_off);
    SOFT_UPLOAD(w->cloth_bufs, s.cl_ptri, ptri);
    SOFT_ALLOC(w->cloth_bufs, s.cl_share, share);
    SOFT_ALLOC(w->cloth_bufs, s.cl_bins, bins);
    SOFT_ALLOC(w->cloth_bufs, s.cl_cols, cols);
    s.n_cloths = n;
    w->cl_rec = rec;
    w->cl_anchor_ids = ids;
    w->cl_anchor_res = res;
    w->cl_max_np = max_np;
    return soft_resolve(w);
}

extern "C" int pxg_get_cloth(pxg_world* w, uint32_t index, pxg_cloth* out) {
    if (!w || !out) return fail(PXG_ERR_INVALID, "pxg_get_cloth: null argument");
    if (index >= w->cl_rec.size()) return fail(PXG_ERR_INVALID, "pxg_get_cloth: no such cloth slot");
    const ClothRec& r = w->cl_rec[index];
    const SoftDev& s = w->sd;
    const std::string at = "pxg_get_cloth: ";
    if ((out->particles && out->n_particles != r.np) || (out->stretch && out->n_stretch != r.n_stretch) ||
        (out->bend && out->n_bend != r.n_bend) || (out->anchors && out->n_anchors != r.n_anchor))
        return fail(PXG_ERR_INVALID, at + "array sizes differ from the uploaded cloth");
    hipStream_t st = w->stream;
    SoftState state;
    HIPCHK(hipMemcpyAsync(&state, s.cl_state + index, sizeof(state), hipMemcpyDeviceToHost, st));
    std::vector<float4> pos(r.np), prev(r.np), vel(r.np), anc(3 * size_t(r.n_anchor));
    std::vector<uint4> con(size_t(r.n_stretch) + r.n_bend);
    if (r.alive && out->particles && r.np) {
        HIPCHK(hipMemcpyAsync(pos.data(), s.cl_pos + r.p0, r.np * sizeof(float4), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(prev.data(), s.cl_prev + r.p0, r.np * sizeof(float4), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(vel.data(), s.cl_vel + r.p0, r.np * sizeof(float4), hipMemcpyDeviceToHost, st));
    }
    if (r.alive && !con.empty() && (out->stretch || out->bend))
        HIPCHK(hipMemcpyAsync(con.data(), s.cl_con + r.c0, con.size() * sizeof(uint4), hipMemcpyDeviceToHost, st));
    if (r.alive && out->anchors && r.n_anchor)
        HIPCHK(hipMemcpyAsync(anc.data(), s.cl_anchor + 3 * size_t(r.a0), anc.size() * sizeof(float4), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (!r.alive) return PXG_OK;
    for (uint32_t i = 0; out->particles && i < r.np; ++i) put_particle(pos[i], prev[i], vel[i], out->particles[i]);
    for (uint32_t i = 0; out->stretch && i < r.n_stretch; ++i) out->stretch[i].lambda = u2f(con[i].w);
    for (uint32_t i = 0; out->bend && i < r.n_bend; ++i) out->bend[i].lambda = u2f(con[r.n_stretch + i].w);
    for (uint32_t a = 0; out->anchors && a < r.n_anchor; ++a) {
        const float4 x = anc[3 * a], y = anc[3 * a + 1];
        pxg_cloth_anchor& o = out->anchors[a];
        o.world_position[0] = x.x; o.world_position[1] = x.y; o.world_position[2] = x.z;
        o.prev_world_position[0] = y.x; o.prev_world_position[1] = y.y; o.prev_world_position[2] = y.z;
    }
    out->awake = state.awake;
    out->enabled = r.enabled;
    out->sleep_timer = state.sleep_timer;
    out->time = state.time;
    for (int k = 0; k < 3; ++k) {
        out->bounds_min[k] = state.bmin[k];
        out->bounds_max[k] = state.bmax[k];
    }
    return PXG_OK;
}

extern "C" int pxg_set_strand_sets(pxg_world* w, const pxg_strand_set* sets, uint32_t n) {
    if (!w || (n && !sets)) return fail(PXG_ERR_INVALID, "pxg_set_strand_sets: null argument");
    if (w->phase != 0) return fail(PXG_ERR_STATE, "pxg_set_strand_sets: not allowed between the two step phases");
    std::vector<StrandRec> rec(n);
    std::vector<SoftState> state(n);
    std::vector<float4> pos, prev, vel, rest_local, rest_offset, strand;
    std::vector<uint4> con;
    std::vector<uint2> batch;
    std::vector<uint32_t> ids, stamp;
    uint32_t tick = 0, max_np = 0;
    for (uint32_t si = 0; si < n; ++si) {
        const pxg_strand_set& g = sets[si];
        StrandRec& r = rec[si];
        std::memset(&r, 0, sizeof(r));
        SoftState& st = state[si];
        std::memset(&st, 0, sizeof(st));
        const std::string at = "pxg_set_strand_sets: set " + std::to_string(si) + ": ";
        r.alive = g.alive && g.n_particles && g.n_strands ? 1u : 0u;
        if (!r.alive) continue;
        const uint32_t np = g.n_particles;
        if (!g.particles || !g.rest_local || !g.rest_offset || !g.strands) return fail(PXG_ERR_INVALID, at + "null array");
        if ((g.n_stretch && !g.stretch) || (g.n_stretch_batches && !g.stretch_batches) ||
            (g.n_collide_bodies && !g.collide_bodies))
            return fail(PXG_ERR_INVALID, at + "null array");
        if (np > kSoftMaxParticles) return fail(PXG_ERR_CAPACITY, at + "more than 8192 particles");
        stamp.assign(np, 0u);
        tick = 0;
        if (const char* e = check_batches(g.stretch_batches, g.n_stretch_batches, g.stretch, g.n_stretch, np, stamp, tick))
            return fail(PXG_ERR_INVALID, at + e);
        ++tick;
        for (uint32_t k = 0; k < g.n_strands; ++k) {  // strands are disjoint ranges of the particles
            const pxg_strand& x = g.strands[k];
            if (!x.count) continue;
            if (x.first >= np || x.count > np - x.first) return fail(PXG_ERR_INVALID, at + "a strand's range is out of range");
            for (uint32_t i = x.first; i < x.first + x.count; ++i) {
                if (stamp[i] == tick) return fail(PXG_ERR_INVALID, at + "two strands share a particle");
                stamp[i] = tick;
            }
        }
        max_np = std::max(max_np, np);
        r.p0 = uint32_t(pos.size());
        r.np = np;
        r.c0 = uint32_t(con.size());
        r.n_stretch = g.n_stretch;
        r.b0 = uint32_t(batch.size());
        r.n_b = g.n_stretch_batches;
        r.s0 = uint32_t(strand.size() / 4);
        r.n_strands = g.n_strands;
        r.cb0 = uint32_t(ids.size() / 2);
        r.n_cb = g.n_collide_bodies;
        r.enabled = g.enabled ? 1u : 0u;
        const pxg_strand_params& q = g.params;
        r.substeps = q.substeps;
        r.iterations = q.iterations;
        r.layer = q.layer;
        r.stretch_compliance = q.stretch_compliance; r.global_shape = q.global_shape;
        r.global_shape_falloff = q.global_shape_falloff; r.local_shape = q.local_shape; r.damping = q.damping;
        r.thickness = q.thickness; r.friction = q.friction; r.gravity_scale = q.gravity_scale;
        r.max_velocity = q.max_velocity; r.wind_turbulence = q.wind_turbulence; r.air_drag = q.air_drag;
        r.sleep_threshold = q.sleep_threshold; r.sleep_time = q.sleep_time;
        for (int k = 0; k < 3; ++k) {
            r.ext[k] = q.external_acceleration[k];
            r.wind[k] = q.wind[k];
        }
        for (uint32_t i = 0; i < np; ++i) {
            const pxg_particle& p = g.particles[i];
            pos.push_back(f4(p.position, p.inv_mass));
            prev.push_back(f4(p.prev_position, p.inv_mass));
            vel.push_back(f4(p.velocity, 0.0f));
            rest_local.push_back(f4(g.rest_local + 3 * size_t(i), 0.0f));
            rest_offset.push_back(f4(g.rest_offset + 3 * size_t(i), 0.0f));
        }
        for (uint32_t i = 0; i < g.n_stretch; ++i)
            con.push_back(make_uint4(g.stretch[i].a, g.stretch[i].b, f2u(g.stretch[i].rest_length), f2u(g.stretch[i].lambda)));
        for (uint32_t k = 0; k < g.n_stretch_batches; ++k)
            batch.push_back(make_uint2(g.stretch_batches[2 * k], g.stretch_batches[2 * k + 1]));
        for (uint32_t k = 0; k < g.n_strands; ++k) {
            const pxg_strand& x = g.strands[k];
            strand.push_back(f4(x.root_position, u2f(x.first)));
            strand.push_back(f4q(x.root_orientation));
            strand.push_back(f4(x.prev_root_position, u2f(x.count)));
            strand.push_back(f4q(x.prev_root_orientation));
        }
        for (uint32_t k = 0; k < g.n_collide_bodies; ++k) {
            ids.push_back(g.collide_bodies[2 * k]);
            ids.push_back(g.collide_bodies[2 * k + 1]);
        }
        st.awake = g.awake ? 1u : 0u;
        st.sleep_timer = g.sleep_timer;
        st.time = g.time;
        particle_bounds(g.particles, np, st.bmin, st.bmax);  // add_strands recomputes them (world.h:638)
    }
    int rr = soft_lds(w, max_np);
    if (rr != PXG_OK) return rr;
    HIPCHK(hipStreamSynchronize(w->stream));
    for (void* p : w->strand_bufs) dfree(w, p);
    w->strand_bufs.clear();
    SoftDev& s = w->sd;
    s.n_sets = 0;
    w->st_rec.clear();
    w->st_collide_ids.clear();
    w->st_collide_res.clear();
    w->st_max_np = 0;
    if (!n) return PXG_OK;
    std::vector<uint32_t> cols(size_t(n) * kSoftMaxColliders), res(ids.size() / 2, kNoBody);
    SoftState* dstate = nullptr;
    SOFT_UPLOAD(w->strand_bufs, s.st, rec);
    SOFT_ALLOC(w->strand_bufs, dstate, state);
    s.st_state = dstate;
    SOFT_ALLOC(w->strand_bufs, s.st_pos, pos);
    SOFT_ALLOC(w->strand_bufs, s.st_prev, prev);
    SOFT_ALLOC(w->strand_bufs, s.st_vel, vel);
    SOFT_ALLOC(w->strand_bufs, s.st_con, con);
    SOFT_UPLOAD(w->strand_bufs, s.st_batch, batch);
    SOFT_UPLOAD(w->strand_bufs, s.st_rest_local, rest_local);
    SOFT_UPLOAD(w->strand_bufs, s.st_rest_offset, rest_offset);
    SOFT_ALLOC(w->strand_bufs, s.st_strand, strand);
    SOFT_UPLOAD(w->strand_bufs, s.st_collide, res);
    SOFT_ALLOC(w->strand_bufs, s.st_cols, cols);
    s.n_sets = n;
    w->st_rec = rec;
    w->st_collide_ids = ids;
    w->st_collide_res = res;
    w->st_max_np = max_np;
    return soft_resolve(w);
}

extern "C" int pxg_get_strand_set(pxg_world* w, uint32_t index, pxg_strand_set* out) {
    if (!w || !out) return fail(PXG_ERR_INVALID, "pxg_get_strand_set: null argument");
    if (index >= w->st_rec.size()) return fail(PXG_ERR_INVALID, "pxg_get_strand_set: no such set slot");
    const StrandRec& r = w->st_rec[index];
    const SoftDev& s = w->sd;
    if ((out->particles && out->n_particles != r.np) || (out->stretch && out->n_stretch != r.n_stretch) ||
        (out->strands && out->n_strands != r.n_strands))
        return fail(PXG_ERR_INVALID, "pxg_get_strand_set: array sizes differ from the uploaded set");
    hipStream_t st = w->stream;
    SoftState state;
    HIPCHK(hipMemcpyAsync(&state, s.st_state + index, sizeof(state), hipMemcpyDeviceToHost, st));
    std::vector<float4> pos(r.np), prev(r.np), vel(r.np), strand(4 * size_t(r.n_strands));
    std::vector<uint4> con(r.n_stretch);
    if (r.alive && out->particles) {
        HIPCHK(hipMemcpyAsync(pos.data(), s.st_pos + r.p0, r.np * sizeof(float4), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(prev.data(), s.st_prev + r.p0, r.np * sizeof(float4), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(vel.data(), s.st_vel + r.p0, r.np * sizeof(float4), hipMemcpyDeviceToHost, st));
    }
    if (r.alive && out->stretch && !con.empty())
        HIPCHK(hipMemcpyAsync(con.data(), s.st_con + r.c0, con.size() * sizeof(uint4), hipMemcpyDeviceToHost, st));
    if (r.alive && out->strands)
        HIPCHK(hipMemcpyAsync(strand.data(), s.st_strand + 4 * size_t(r.s0), strand.size() * sizeof(float4),
                              hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (!r.alive) return PXG_OK;
    for (uint32_t i = 0; out->particles && i < r.np; ++i) put_particle(pos[i], prev[i], vel[i], out->particles[i]);
    for (uint32_t i = 0; out->stretch && i < r.n_stretch; ++i) out->stretch[i].lambda = u2f(con[i].w);
    for (uint32_t k = 0; out->strands && k < r.n_strands; ++k) {
        pxg_strand& o = out->strands[k];
        const float4 a = strand[4 * k], q = strand[4 * k + 1], b = strand[4 * k + 2], pq = strand[4 * k + 3];
        o.root_position[0] = a.x; o.root_position[1] = a.y; o.root_position[2] = a.z;
        o.root_orientation[0] = q.x; o.root_orientation[1] = q.y; o.root_orientation[2] = q.z; o.root_orientation[3] = q.w;
        o.prev_root_position[0] = b.x; o.prev_root_position[1] = b.y; o.prev_root_position[2] = b.z;
        o.prev_root_orientation[0] = pq.x; o.prev_root_orientation[1] = pq.y; o.prev_root_orientation[2] = pq.z;
        o.prev_root_orientation[3] = pq.w;
    }
    out->awake = state.awake;
    out->enabled = r.enabled;
    out->sleep_timer = state.sleep_timer;
    out->time = state.time;
    for (int k = 0; k < 3; ++k) {
        out->bounds_min[k] = state.bmin[k];
        out->bounds_max[k] = state.bmax[k];
    }
    return PXG_OK;
}

extern "C" int pxg_set_strand_roots(pxg_world* w, uint32_t index, const float* positions, const float* orientations,
                                    uint32_t n) {
    if (!w || (n && !positions)) return fail(PXG_ERR_INVALID, "pxg_set_strand_roots: null argument");
    if (index >= w->st_rec.size()) return fail(PXG_ERR_INVALID, "pxg_set_strand_roots: no such set slot");
    const StrandRec& r = w->st_rec[index];
    const uint32_t m = std::min(n, r.n_strands);
    if (!r.alive || !m) return PXG_OK;
    float4* dev = w->sd.st_strand + 4 * size_t(r.s0);
    std::vector<float4> strand(4 * size_t(m));
    hipStream_t st = w->stream;
    HIPCHK(hipMemcpyAsync(strand.data(), dev, strand.size() * sizeof(float4), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (uint32_t k = 0; k < m; ++k) {  // prev = current; current = the new root
        float4* x = &strand[4 * k];
        x[2] = make_float4(x[0].x, x[0].y, x[0].z, x[2].w);
        x[3] = x[1];
        x[0] = make_float4(positions[3 * k], positions[3 * k + 1], positions[3 * k + 2], x[0].w);
        if (orientations) x[1] = f4q(orientations + 4 * size_t(k));
    }
    HIPCHK(hipMemcpyAsync(dev, strand.data(), strand.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return PXG_OK;
}

// the once-per-frame cloth and hair pass of step() (world.h:1241-1256)
static int step_soft(pxg_world* w, float frame, uint32_t substeps_run) {
    SoftDev& s = w->sd;
    if ((!s.n_cloths && !s.n_sets) || !(frame > 0.0f)) return PXG_OK;
    const pxg_config& c = w->cfg;
    if (w->soft_resolve) {
        int r = soft_resolve(w);
        if (r != PXG_OK) return r;
    }
    hipStream_t st = w->stream;
    if (substeps_run == 0 && w->shapes_ready) launch_refit(w->d, frame, c.aabb_margin, st);  // sync_broadphase_
    // world.h:1753-1757: substeps of fixed_dt / cloth_substeps in this frame, 1 to 16
    const float internal = std::max(c.fixed_dt / float(std::max(1u, c.cloth_substeps)), 1e-6f);
    const uint32_t fs = uint32_t(std::clamp<long>(std::lround(frame / internal), 1, 16));
    pxd::v3 g;
    g.x = c.gravity[0];
    g.y = c.gravity[1];
    g.z = c.gravity[2];
    launch_cloths(w->d, s, w->cl_max_np, frame, fs, c.cloth_iterations, g, c.aabb_margin, st);
    launch_strands(w->d, s, w->st_max_np, frame, fs, c.cloth_iterations, g, c.aabb_margin, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PXG_ERR_HIP, std::string("cloth / hair launch: ") + hipGetErrorString(e));
    std::vector<SoftState> cs(s.n_cloths), ss(s.n_sets);
    if (s.n_cloths) HIPCHK(hipMemcpyAsync(cs.data(), s.cl_state, cs.size() * sizeof(SoftState), hipMemcpyDeviceToHost, st));
    if (s.n_sets) HIPCHK(hipMemcpyAsync(ss.data(), s.st_state, ss.size() * sizeof(SoftState), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (const SoftState& x : cs)
        if (x.error) return fail(PXG_ERR_CAPACITY, "pxg_step: a cloth met more than 256 colliders in a frame");
    for (const SoftState& x : ss)
        if (x.error) return fail(PXG_ERR_CAPACITY, "pxg_step: a hair groom met more than 256 colliders in a frame");
    return PXG_OK;
}

extern "C" int pxg_state_hash(pxg_world* w, uint64_t* out) {
    if (!w || !out) return fail(PXG_ERR_INVALID, "pxg_state_hash: null argument");
    uint32_t n = w->n_bodies;
    std::vector<pxg_body> b(n);
    int r = pxg_get_bodies(w, b.data(), n);
    if (r != PXG_OK) return r;
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void* p, size_t len) {
        const uint8_t* c = static_cast<const uint8_t*>(p);
        for (size_t i = 0; i < len; ++i) { h ^= c[i]; h *= 1099511628211ull; }
    };
    for (uint32_t i = 0; i < n; ++i) {
        if (!b[i].alive) continue;
        mix(b[i].position, 12);
        mix(b[i].orientation, 16);
        mix(b[i].linear_velocity, 12);
        mix(b[i].angular_velocity, 12);
    }
    for (const ClothRec& r : w->cl_rec) {  // alive cloths' particles (world.h:1352-1357)
        if (!r.alive) continue;
        std::vector<float4> pos(r.np), vel(r.np);
        HIPCHK(hipMemcpyAsync(pos.data(), w->sd.cl_pos + r.p0, r.np * sizeof(float4), hipMemcpyDeviceToHost, w->stream));
        HIPCHK(hipMemcpyAsync(vel.data(), w->sd.cl_vel + r.p0, r.np * sizeof(float4), hipMemcpyDeviceToHost, w->stream));
        HIPCHK(hipStreamSynchronize(w->stream));
        for (uint32_t i = 0; i < r.np; ++i) {
            mix(&pos[i], 12);
            mix(&vel[i], 12);
        }
    }
    *out = h;
    return PXG_OK;
}

namespace pxg { void mlp_to_device_image(const pxg_mlp* net, MlpDev* img); }

static int ensure_learn_buffers(pxg_world* w) {
    if (w->d.feat) return PXG_OK;
    if (!w->cap_m) return fail(PXG_ERR_STATE, "upload shapes before enabling the learned warm start");
    DevState& d = w->d;
    DALLOC(d.feat, size_t(4) * w->cap_m * kMlpIn);
    DALLOC(d.miss_idx, size_t(4) * w->cap_m);
    DALLOC(d.pred, size_t(4) * w->cap_m * kMlpOut);
    return PXG_OK;
}

extern "C" int pxg_set_mlp(pxg_world* w, const pxg_mlp* net) {
    if (!w) return fail(PXG_ERR_INVALID, "pxg_set_mlp: null handle");
    if (!net) {
        w->d.mlp = nullptr;
        return PXG_OK;
    }
    int r = ensure_learn_buffers(w);
    if (r != PXG_OK) return r;
    if (!w->d_mlp) {
        MlpDev* p = nullptr;
        DALLOC(p, 1);
        w->d_mlp = p;
    }
    MlpDev img;
    mlp_to_device_image(net, &img);
    HIPCHK(hipMemcpyAsync(w->d_mlp, &img, sizeof(img), hipMemcpyHostToDevice, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    w->d.mlp = w->d_mlp;
    return PXG_OK;
}

extern "C" int pxg_set_recording(pxg_world* w, uint32_t on) {
    if (!w) return fail(PXG_ERR_INVALID, "pxg_set_recording: null handle");
    if (on) {
        int r = ensure_learn_buffers(w);
        if (r != PXG_OK) return r;
    }
    w->d.record = on ? 1u : 0u;
    return PXG_OK;
}

extern "C" int pxg_get_records(pxg_world* w, float* features, float* impulses, uint32_t* missed, uint32_t cap, uint32_t* n) {
    if (!w || !n) return fail(PXG_ERR_INVALID, "pxg_get_records: null argument");
    {
        int fr = finish_results(w);
        if (fr != PXG_OK) return fr;
    }
    *n = 0;
    if (!w->d.record || !w->d.feat) return fail(PXG_ERR_STATE, "pxg_get_records: recording is off");
    uint32_t A = w->last_active;
    if (!A) return PXG_OK;
    DevState& d = w->d;
    hipStream_t st = w->stream;
    std::vector<uint32_t> cnt(A);
    std::vector<float> f(size_t(4) * A * kMlpIn);
    std::vector<float4> imp(4 * size_t(A)), aux(4 * size_t(A));
    HIPCHK(hipMemcpyAsync(cnt.data(), d.sm_cnt, A * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(f.data(), d.feat, f.size() * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(imp.data(), d.sp_imp, imp.size() * sizeof(float4), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(aux.data(), d.sp_aux, aux.size() * sizeof(float4), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    uint32_t k = 0;
    for (uint32_t s = 0; s < A; ++s) {
        for (uint32_t t = 0; t < cnt[s]; ++t, ++k) {
            if (k >= cap) continue;
            size_t pidx = 4 * size_t(s) + t;
            if (features) std::memcpy(features + size_t(k) * kMlpIn, &f[pidx * kMlpIn], kMlpIn * 4);
            if (impulses) { impulses[3 * k] = imp[pidx].x; impulses[3 * k + 1] = imp[pidx].y; impulses[3 * k + 2] = imp[pidx].z; }
            if (missed) missed[k] = aux[pidx].w != 0.0f ? 1u : 0u;
        }
    }
    *n = k;
    return PXG_OK;
}

extern "C" int pxg_get_predictions(pxg_world* w, uint64_t* keys, float* imp3, uint32_t cap, uint32_t* n) {
    if (!w || !n) return fail(PXG_ERR_INVALID, "pxg_get_predictions: null argument");
    {
        int fr = finish_results(w);
        if (fr != PXG_OK) return fr;
    }
    *n = 0;
    if (!w->d.mlp || w->d.record) return PXG_OK;
    DevState& d = w->d;
    hipStream_t st = w->stream;
    HIPCHK(hipMemcpyAsync(w->h_cnt, d.cnt, sizeof(Counters), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    uint32_t P = w->h_cnt->predicted, A = w->last_active;
    if (!P) return PXG_OK;
    std::vector<uint32_t> mi(P);
    std::vector<float> pr(3 * size_t(P));
    std::vector<uint2> ab(A);
    std::vector<float4> aux(4 * size_t(A));
    HIPCHK(hipMemcpyAsync(mi.data(), d.miss_idx, P * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(pr.data(), d.pred, pr.size() * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(ab.data(), d.sm_ab, A * sizeof(uint2), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(aux.data(), d.sp_aux, aux.size() * sizeof(float4), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<std::pair<uint64_t, uint32_t>> e;
    for (uint32_t j = 0; j < P; ++j) {
        uint32_t p = mi[j], s = p / 4;
        uint32_t fid;
        std::memcpy(&fid, &aux[p].z, 4);
        uint32_t lo = std::min(ab[s].x, ab[s].y), hi = std::max(ab[s].x, ab[s].y);
        uint64_t key = (uint64_t(lo) << 40) ^ (uint64_t(hi) << 12) ^ uint64_t(fid);
        e.push_back({key, j});
    }
    std::sort(e.begin(), e.end());
    *n = P;
    for (uint32_t i = 0; i < P && i < cap; ++i) {
        if (keys) keys[i] = e[i].first;
        if (imp3) for (int k = 0; k < 3; ++k) imp3[3 * i + k] = pr[3 * size_t(e[i].second) + k];
    }
    return PXG_OK;
}

// Profiling of the dominant kernel (the velocity pass): HIP events on the world's stream around
// every launch. No counterpart in the reference.
extern "C" int pxg_set_profiling(pxg_world* w, uint32_t on) {
    if (!w) return fail(PXG_ERR_INVALID, "pxg_set_profiling: null handle");
    w->profiling = on != 0;
    return PXG_OK;
}

extern "C" int pxg_get_profile(pxg_world* w, pxg_profile* out) {
    if (!w || !out) return fail(PXG_ERR_INVALID, "pxg_get_profile: null argument");
    HIPCHK(hipStreamSynchronize(w->stream));
    for (auto& pr : w->ev_pending) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
        w->prof_vel_ms += ms;
        w->prof_vel_launches += 1;
        w->ev_pool.push_back(pr.first);
        w->ev_pool.push_back(pr.second);
    }
    w->ev_pending.clear();
    HIPCHK(hipMemcpy(w->h_cnt, w->d.cnt, sizeof(Counters), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    out->velocity_ms = w->prof_vel_ms;
    out->velocity_launches = w->prof_vel_launches;
    out->velocity_bytes = double(w->h_cnt->prof_vel_bytes);
    out->tiled_colour_ticks = w->h_cnt->prof_stage[0];
    out->tiled_stage_ticks = w->h_cnt->prof_stage[1];
    out->tiled_barrier_ticks = w->h_cnt->prof_stage[2];
    out->tiled_residual_ticks = w->h_cnt->prof_stage[3];
    w->prof_vel_ms = 0.0;
    w->prof_vel_launches = 0;
    unsigned long long zero[5] = {0, 0, 0, 0, 0};
    HIPCHK(hipMemcpy(&w->d.cnt->prof_vel_bytes, &zero[0], sizeof(zero[0]), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(&w->d.cnt->prof_stage[0], &zero[1], 4 * sizeof(zero[0]), hipMemcpyHostToDevice));
    return PXG_OK;
}
