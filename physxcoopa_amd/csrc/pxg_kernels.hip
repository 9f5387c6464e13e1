// Step kernels of one world on gfx950 (CDNA4): integration, broadphase (x-y grid over fat
// AABBs plus a brute-force path for large shapes), narrowphase, manifold compaction,
// contact-graph colouring, the colour-by-colour solver passes, the warm-start hash table,
// islands/sleep and velocity integration. Launch shapes: 256-thread blocks (4 wave64s), one
// work item per thread, grid-stride where the item count lives on the device.
#include "pxg_state.h"
#include "pxg_device.h"
#include "pxg_mesh.h"
#include "pxg_joints.h"
#include "pxg_mlp.h"
#include "pxg_shapes.h"

using namespace pxd;

namespace pxg {

namespace {

constexpr unsigned long long kEmptyKey = ~0ull;

__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(u);
}
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t hash64(unsigned long long k) {
    return hash32(uint32_t(k) ^ hash32(uint32_t(k >> 32) + 0x9e3779b9u));
}

__device__ __forceinline__ bool body_active(uint32_t f) {  // dynamic and awake (solver.h:86-95)
    return (f & BF_TYPE) == 2u && (f & BF_AWAKE) && (f & BF_ALIVE);
}

__device__ __forceinline__ m3 load_iiw(const float4* p, uint32_t i) {
    float4 a = p[3 * i], b = p[3 * i + 1], c = p[3 * i + 2];
    m3 m;
    m.c0 = mk(a.x, a.y, a.z);
    m.c1 = mk(a.w, b.x, b.y);
    m.c2 = mk(b.z, b.w, c.x);
    return m;
}
__device__ __forceinline__ void store_iiw(float4* p, uint32_t i, const m3& m, float tail0 = 0.0f, float tail1 = 0.0f,
                                          float tail2 = 0.0f) {
    p[3 * i] = make_float4(m.c0.x, m.c0.y, m.c0.z, m.c1.x);
    p[3 * i + 1] = make_float4(m.c1.y, m.c1.z, m.c2.x, m.c2.y);
    p[3 * i + 2] = make_float4(m.c2.z, tail0, tail1, tail2);
}

__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kBlock / 64];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wid) off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    if (total) *total = tot;
    return off + x - v;
}

// Wave-wide reductions; every lane of the wave must reach the call.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { uint32_t u = __shfl_xor(v, o, 64); v = u < v ? u : v; }
    return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { uint32_t u = __shfl_xor(v, o, 64); v = u > v ? u : v; }
    return v;
}
__device__ __forceinline__ bool lane0() { return (threadIdx.x & 63) == 0; }

// ------------------------------------------------------------------------------------------
// generic helpers
// ------------------------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, size_t n) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

// 3-phase exclusive scan: each block covers 1024 elements (4 per thread).
__global__ void k_scan_reduce(const uint32_t* in, uint32_t n, uint32_t* bsum) {
    uint32_t base = blockIdx.x * 1024u + threadIdx.x * 4u;
    uint32_t s = 0;
    for (int k = 0; k < 4; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
__global__ void k_scan_bsums(uint32_t* bsum, uint32_t nb, uint32_t* total) {
    // single block of 256 threads, up to 1024 block sums (4 per thread)
    uint32_t base = threadIdx.x * 4u;
    uint32_t v[4], s = 0;
    for (int k = 0; k < 4; ++k) { v[k] = (base + k < nb) ? bsum[base + k] : 0u; s += v[k]; }
    uint32_t tot;
    uint32_t off = block_exclusive_scan(s, &tot);
    for (int k = 0; k < 4; ++k) {
        if (base + k < nb) bsum[base + k] = off;
        off += v[k];
    }
    if (threadIdx.x == 0 && total) *total = tot;
}
__global__ void k_scan_down(const uint32_t* in, uint32_t* out, uint32_t n, const uint32_t* bsum) {
    uint32_t base = blockIdx.x * 1024u + threadIdx.x * 4u;
    uint32_t v[4], s = 0;
    for (int k = 0; k < 4; ++k) { v[k] = (base + k < n) ? in[base + k] : 0u; s += v[k]; }
    uint32_t off = block_exclusive_scan(s, nullptr) + bsum[blockIdx.x];
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = off;
        off += v[k];
    }
}

// ------------------------------------------------------------------------------------------
// bodies
// ------------------------------------------------------------------------------------------
// update_world_inertia for every body (integrator.h:57-68), as add_body() does (world.h:113)
__global__ void k_init_bodies(DevState d) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nb) return;
    uint32_t f = d.b_flags[i];
    m3 m;
    if ((f & BF_TYPE) != 2u) { m.c0 = splat(0.0f); m.c1 = splat(0.0f); m.c2 = splat(0.0f); }
    else m = world_inertia(ldq(d.b_rot[i]), ld3(d.b_iil[i]));
    store_iiw(d.b_iiw, i, m);
}

// Writes n staged body records (9 float4 fields each, SoA over the n records, then the flag words)
// into bodies idx[k]; with `inertia` the written bodies' world inverse inertia is recomputed as
// k_init_bodies does for all of them
__global__ void k_scatter_bodies(DevState d, const uint32_t* idx, const float4* rec, const uint32_t* flags, uint32_t n,
                                 int inertia) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = idx[k];
    d.b_pos[i] = rec[k];
    d.b_rot[i] = rec[n + k];
    d.b_vel[i] = rec[2 * n + k];
    d.b_angv[i] = rec[3 * n + k];
    d.b_force[i] = rec[4 * n + k];
    d.b_torque[i] = rec[5 * n + k];
    d.b_iil[i] = rec[6 * n + k];
    d.b_prev_pos[i] = rec[7 * n + k];
    d.b_prev_rot[i] = rec[8 * n + k];
    const uint32_t f = flags[k];
    d.b_flags[i] = f;
    if (inertia) {
        m3 m;
        if ((f & BF_TYPE) != 2u) { m.c0 = splat(0.0f); m.c1 = splat(0.0f); m.c2 = splat(0.0f); }
        else m = world_inertia(ldq(rec[n + k]), ld3(rec[6 * n + k]));
        store_iiw(d.b_iiw, i, m);
    }
}

// apply_gravity + integrate_forces (world.h:1283-1295, integrator.h:31-35, 78-94)
__global__ void k_integrate_forces(DevState d, StepParams p) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {  // reset this step's counters
        Counters* c = d.cnt;
        c->pairs = 0; c->manifolds = 0; c->points = 0; c->active = 0; c->ordered = 0; c->remaining = 0;
        c->colours = 0; c->n_large = 0; c->awake = 0; c->predicted = 0; c->residual_bits = 0;
        for (int k = 0; k < 3; ++k) c->red[k] = 0xFFFFFFFFu;
        for (int k = 3; k < 12; ++k) c->red[k] = 0u;
        c->passes = 0; c->res_slot[0] = 0u; c->res_slot[1] = 0u;
    }
    if (i >= d.nb) return;
    uint32_t f = d.b_flags[i];
    if (!(f & BF_ALIVE)) return;
    bool dyn = (f & BF_TYPE) == 2u, awake = (f & BF_AWAKE) != 0;
    float4 fo = d.b_force[i], to = d.b_torque[i];
    if (!dyn || !awake) {
        d.b_force[i] = make_float4(0.0f, 0.0f, 0.0f, fo.w);
        d.b_torque[i] = make_float4(0.0f, 0.0f, 0.0f, to.w);
        return;
    }
    float4 vm = d.b_vel[i], wm = d.b_angv[i];
    v3 force = ld3(fo), torque = ld3(to);
    if (f & BF_GRAVITY) force = force + mk(p.gx, p.gy, p.gz) * wm.w;  // force_accum += g * mass
    v3 v = ld3(vm), w = ld3(wm);
    v = v + (force * vm.w) * p.h;
    m3 I = load_iiw(d.b_iiw, i);
    w = w + mul(I, torque) * p.h;
    v = v * (1.0f / (1.0f + p.h * fo.w));
    w = w * (1.0f / (1.0f + p.h * to.w));
    uint32_t cons = f >> BF_CONSTRAINT_SHIFT;
    if (cons & 1u) v.x = 0.0f;
    if (cons & 2u) v.y = 0.0f;
    if (cons & 4u) v.z = 0.0f;
    if (cons & 8u) w.x = 0.0f;
    if (cons & 16u) w.y = 0.0f;
    if (cons & 32u) w.z = 0.0f;
    d.b_vel[i] = st3(v, vm.w);
    d.b_angv[i] = st3(w, wm.w);
    d.b_force[i] = make_float4(0.0f, 0.0f, 0.0f, fo.w);
    d.b_torque[i] = make_float4(0.0f, 0.0f, 0.0f, to.w);
}

// ------------------------------------------------------------------------------------------
// broadphase
// ------------------------------------------------------------------------------------------
// create_shape_proxy_ (world.h:1369-1382): fat = tight.expand(margin) at the owner's pose
__global__ void k_init_proxies(DevState d, StepParams p) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.ns) return;
    uint4 info = d.s_info[i];
    if (!shape_live(d, info)) {
        d.s_fmin[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        d.s_fmax[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    ShapeGeom s = load_shape(d, i, info);
    v3 mn, mx;
    shape_bounds(d, s, ld3(d.b_pos[info.x]), ldq(d.b_rot[info.x]), mn, mx);
    d.s_fmin[i] = st3(mn - splat(p.margin), 0.0f);
    d.s_fmax[i] = st3(mx + splat(p.margin), 0.0f);
}

// compute_broadphase_bounds_ + move_proxy (world.h:1489-1518, aabb_tree.h:76-85); large-shape
// list; reductions for the grid over small shapes' fat AABBs.
__global__ void k_bounds(DevState d, StepParams p) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // min x, y, z and max x, y, z of fat min corners; max extents x, y, z
    uint32_t rmn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, rmx[3] = {0u, 0u, 0u}, rex[3] = {0u, 0u, 0u};
    if (i < d.ns) {
        uint4 info = d.s_info[i];
        if (!shape_live(d, info)) {
            d.s_tmin[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        } else {
            ShapeGeom s = load_shape(d, i, info);
            v3 mn, mx;
            uint32_t ob = info.x;
            shape_bounds(d, s, ld3(d.b_pos[ob]), ldq(d.b_rot[ob]), mn, mx);
            v3 fmn = ld3(d.s_fmin[i]), fmx = ld3(d.s_fmax[i]);
            if (!(info.y & SF_STATIC)) {
                v3 disp = ld3(d.b_vel[ob]) * p.h;
                v3 smn = vmin(mn, mn + disp), smx = vmax(mx, mx + disp);
                bool contains = smn.x >= fmn.x && smn.x <= fmx.x && smn.y >= fmn.y && smn.y <= fmx.y &&
                                smn.z >= fmn.z && smn.z <= fmx.z && smx.x >= fmn.x && smx.x <= fmx.x &&
                                smx.y >= fmn.y && smx.y <= fmx.y && smx.z >= fmn.z && smx.z <= fmx.z;
                if (!contains) {
                    fmn = smn - splat(p.margin);
                    fmx = smx + splat(p.margin);
                    d.s_fmin[i] = st3(fmn, 0.0f);
                    d.s_fmax[i] = st3(fmx, 0.0f);
                }
            }
            float ex = fmx.x - fmn.x, ey = fmx.y - fmn.y, ez = fmx.z - fmn.z;
            bool large = ex > p.large_extent || ey > p.large_extent || ez > p.large_extent;
            d.s_tmin[i] = st3(mn, large ? 1.0f : 0.0f);
            d.s_tmax[i] = st3(mx, 0.0f);
            if (large) {
                uint32_t k = atomicAdd(&d.cnt->n_large, 1u);
                if (k < uint32_t(kMaxLarge)) d.large_list[k] = i;
                else atomicOr(&d.cnt->err, ERR_LARGE);
            } else {
                rmn[0] = rmx[0] = f2o(fmn.x);
                rmn[1] = rmx[1] = f2o(fmn.y);
                rmn[2] = rmx[2] = f2o(fmn.z);
                rex[0] = f2o(ex); rex[1] = f2o(ey); rex[2] = f2o(ez);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        rmn[k] = wave_min(rmn[k]);
        rmx[k] = wave_max(rmx[k]);
        rex[k] = wave_max(rex[k]);
    }
    if (lane0() && rmx[0] != 0u) {
        Counters* c = d.cnt;
        for (int k = 0; k < 3; ++k) {
            atomicMin(&c->red[k], rmn[k]);
            atomicMax(&c->red[3 + k], rmx[k]);
            atomicMax(&c->red[6 + k], rex[k]);
        }
    }
}

// 3-D uniform grid over the small shapes' fat-min corners. The cell is at least the largest fat
// extent, so every candidate partner's min corner lies in a bounded block of cells.
__global__ void k_grid_setup(DevState d) {
    Counters* c = d.cnt;
    float o[3] = {0.0f, 0.0f, 0.0f}, ext[3] = {0.0f, 0.0f, 0.0f};
    int n[3] = {1, 1, 1};
    float cell = 1.0f;
    if (c->red[0] != 0xFFFFFFFFu) {
        float mx[3];
        for (int k = 0; k < 3; ++k) {
            o[k] = o2f(c->red[k]);
            mx[k] = o2f(c->red[3 + k]);
            ext[k] = o2f(c->red[6 + k]);
        }
        cell = fmaxf(fmaxf(fmaxf(ext[0], ext[1]), ext[2]), 1e-3f);
        for (int it = 0; it < 64; ++it) {
            float f[3];
            for (int k = 0; k < 3; ++k) f[k] = floorf((mx[k] - o[k]) / cell) + 1.0f;
            if (f[0] * f[1] * f[2] <= float(kMaxCells)) {
                for (int k = 0; k < 3; ++k) n[k] = int(f[k]);
                break;
            }
            cell *= 2.0f;
        }
    }
    c->grid[0] = o[0]; c->grid[1] = o[1]; c->grid[2] = o[2]; c->grid[3] = 1.0f / cell;
    c->grid[4] = ext[1]; c->grid[5] = ext[2];
    c->grid[6] = float(n[0]); c->grid[7] = float(n[1]); c->grid[8] = float(n[2]);
}

struct Grid {
    float o[3], inv, exty, extz;
    int n[3];
};
__device__ __forceinline__ Grid load_grid(const Counters* c) {
    Grid g;
    g.o[0] = c->grid[0]; g.o[1] = c->grid[1]; g.o[2] = c->grid[2]; g.inv = c->grid[3];
    g.exty = c->grid[4]; g.extz = c->grid[5];
    g.n[0] = int(c->grid[6]); g.n[1] = int(c->grid[7]); g.n[2] = int(c->grid[8]);
    return g;
}
__device__ __forceinline__ int cell_of(const Grid& g, int k, float x) {
    float f = floorf((x - g.o[k]) * g.inv);
    return f < 0.0f ? 0 : (f > float(g.n[k] - 1) ? g.n[k] - 1 : int(f));
}
__device__ __forceinline__ int cell_index(const Grid& g, v3 p) {
    return (cell_of(g, 2, p.z) * g.n[1] + cell_of(g, 1, p.y)) * g.n[0] + cell_of(g, 0, p.x);
}

__global__ void k_grid_clear(DevState d) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kMaxCells; i += gridDim.x * blockDim.x) {
        d.cell_count[i] = 0u;
        d.cell_cursor[i] = 0u;
    }
}
__global__ void k_grid_count(DevState d) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.ns) return;
    if (d.s_tmin[i].w != 0.0f) return;  // inactive (-1) or large (1)
    Grid g = load_grid(d.cnt);
    atomicAdd(&d.cell_count[cell_index(g, ld3(d.s_fmin[i]))], 1u);
}
__global__ void k_grid_scatter(DevState d) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.ns) return;
    float4 t = d.s_tmin[i];
    if (t.w != 0.0f) return;
    Grid g = load_grid(d.cnt);
    float4 f = d.s_fmin[i];
    int c = cell_index(g, ld3(f));
    uint32_t slot = d.cell_start[c] + atomicAdd(&d.cell_cursor[c], 1u);
    uint4 info = d.s_info[i];
    d.ci_fmin[slot] = make_float4(f.x, f.y, f.z, __uint_as_float(i));
    float4 fx = d.s_fmax[i];
    d.ci_fmax[slot] = make_float4(fx.x, fx.y, fx.z, __uint_as_float(info.x));
    d.ci_tmin[slot] = make_float4(t.x, t.y, t.z, (info.y & SF_STATIC) ? 1.0f : 0.0f);
    d.ci_tmax[slot] = d.s_tmax[i];
}

// The reference pair rule (world.h:1545-1577 with aabb.h:35-39): owners differ, not both static,
// and tight(i) overlaps fat(j) or tight(j) overlaps fat(i).
__device__ __forceinline__ bool ovl(v3 amn, v3 amx, v3 bmn, v3 bmx) {
    return amn.x <= bmx.x && amx.x >= bmn.x && amn.y <= bmx.y && amx.y >= bmn.y && amn.z <= bmx.z && amx.z >= bmn.z;
}

// One emitter per pair: small-small pairs by the shape with the smaller (fat min x, index),
// small-large pairs by the small shape, large-large pairs by the smaller index. Each emitter's
// candidate cells are rows of contiguous cells (one row per (z, y) cell pair of its fat box);
// kPairLanes threads share an emitter, thread r taking rows r, r + kPairLanes, ... (thread 0 also
// takes the large-shape list), so the scan is spread over many more lanes than emitters. Mode 0
// counts per (emitter, lane); mode 1 writes at the scanned offsets; the emitter's segment is then
// sorted by (a, b) (k_pairs_sort), so the pair list does not depend on the lane split.
template <int MODE>
__global__ void k_pairs(DevState d) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = gid / kPairLanes, r = gid % kPairLanes;
    if (i >= d.ns) return;
    float4 ti4 = d.s_tmin[i];
    if (ti4.w < 0.0f) {
        if (MODE == 0) d.pair_count[gid] = 0u;
        return;
    }
    uint4 info = d.s_info[i];
    uint32_t own_i = info.x;
    bool st_i = (info.y & SF_STATIC) != 0;
    bool large_i = ti4.w > 0.0f;
    v3 tmn = ld3(ti4), tmx = ld3(d.s_tmax[i]);
    v3 fmn = ld3(d.s_fmin[i]), fmx = ld3(d.s_fmax[i]);
    uint32_t n = 0, off = 0, cap = d.cap_pairs;
    if (MODE == 1) off = d.pair_off[gid];
    const Counters* c = d.cnt;

    auto emit = [&](uint32_t j) {
        if (MODE == 1) {
            uint32_t k = off + n;
            if (k < cap) d.pairs[k] = make_uint2(i < j ? i : j, i < j ? j : i);
        }
        ++n;
    };

    if (!large_i) {
        Grid g = load_grid(c);
        int x0 = cell_of(g, 0, fmn.x), x1 = cell_of(g, 0, fmx.x);
        int y0 = cell_of(g, 1, fmn.y - g.exty), y1 = cell_of(g, 1, fmx.y);
        int z0 = cell_of(g, 2, fmn.z - g.extz), z1 = cell_of(g, 2, fmx.z);
        const uint32_t ny = uint32_t(y1 - y0 + 1), rows = uint32_t(z1 - z0 + 1) * ny;
        for (uint32_t q = r; q < rows; q += kPairLanes) {
            int cz = z0 + int(q / ny), cy = y0 + int(q % ny);
            int row = (cz * g.n[1] + cy) * g.n[0];
            // cells x0..x1 of one row are contiguous in the cell-ordered arrays
            uint32_t s0 = d.cell_start[row + x0];
            uint32_t s1 = d.cell_start[row + x1] + d.cell_count[row + x1];
            for (uint32_t sl = s0; sl < s1; ++sl) {
                float4 jf = d.ci_fmin[sl];
                uint32_t j = __float_as_uint(jf.w);
                if (!(jf.x > fmn.x || (jf.x == fmn.x && j > i))) continue;
                if (jf.x > fmx.x) continue;
                float4 jF = d.ci_fmax[sl];
                if (__float_as_uint(jF.w) == own_i) continue;
                float4 jt = d.ci_tmin[sl];
                if (st_i && jt.w != 0.0f) continue;
                float4 jT = d.ci_tmax[sl];
                if (ovl(tmn, tmx, ld3(jf), ld3(jF)) || ovl(ld3(jt), ld3(jT), fmn, fmx)) emit(j);
            }
        }
    }
    if (r == 0) {
        uint32_t nl = c->n_large < uint32_t(kMaxLarge) ? c->n_large : uint32_t(kMaxLarge);
        for (uint32_t k = 0; k < nl; ++k) {
            uint32_t j = d.large_list[k];
            if (j == i) continue;
            if (large_i && j < i) continue;
            uint4 ij = d.s_info[j];
            if (ij.x == own_i) continue;
            if (st_i && (ij.y & SF_STATIC)) continue;
            if (ovl(tmn, tmx, ld3(d.s_fmin[j]), ld3(d.s_fmax[j])) || ovl(ld3(d.s_tmin[j]), ld3(d.s_tmax[j]), fmn, fmx)) emit(j);
        }
    }
    if (MODE == 0) d.pair_count[gid] = n;
}
// insertion sort of each emitter's segment by (a, b)
__global__ void k_pairs_sort(DevState d) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.ns) return;
    uint32_t b0 = d.pair_off[i * kPairLanes];
    uint32_t b1 = (i + 1 < d.ns) ? d.pair_off[(i + 1) * kPairLanes] : d.cnt->pairs;
    if (b1 > d.cap_pairs || b1 - b0 < 2u) return;
    uint2* seg = d.pairs + b0;
    uint32_t n = b1 - b0;
    for (uint32_t a = 1; a < n; ++a) {
        uint2 v = seg[a];
        unsigned long long kv = (static_cast<unsigned long long>(v.x) << 32) | v.y;
        uint32_t b = a;
        while (b > 0) {
            uint2 u = seg[b - 1];
            unsigned long long ku = (static_cast<unsigned long long>(u.x) << 32) | u.y;
            if (ku <= kv) break;
            seg[b] = u;
            --b;
        }
        seg[b] = v;
    }
}

// ------------------------------------------------------------------------------------------
// narrowphase (world.h:1598-1654, collision/narrowphase.h:175-335)
// ------------------------------------------------------------------------------------------
// joint_blocks_collision_ and ignored_pairs_ (world.h:1619-1620): binary search of the sorted
// body-pair keys
__device__ __forceinline__ bool contacts_off(const DevState& d, uint32_t a, uint32_t b) {
    if (d.n_nc == 0u) return false;
    const unsigned long long key = (static_cast<unsigned long long>(a < b ? a : b) << 32) | (a < b ? b : a);
    uint32_t lo = 0, hi = d.n_nc;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (d.nc_keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo < d.n_nc && d.nc_keys[lo] == key;
}
__device__ __forceinline__ void store_pair_manifold(const DevState& d, uint32_t k, uint32_t i, uint32_t j, uint4 ii,
                                                    uint4 ij, const Man& m) {
    const float4 mi = d.s_mat[i], mj = d.s_mat[j];
    const float fr = combine(mi.y, mj.y, int(ii.w & 255u), int(ij.w & 255u));
    const float re = combine(mi.z, mj.z, int(ii.w >> 8), int(ij.w >> 8));
    const uint32_t trig = ((ii.y | ij.y) & SF_TRIGGER) ? 1u : 0u;
    d.pm_ab[k] = make_uint2(ii.x, ij.x);
    d.pm_nf[k] = st3(m.n, fr);
    d.pm_meta[k] = make_uint4(uint32_t(m.count) | (trig << 8), __float_as_uint(re), 0u, 0u);
    for (int t = 0; t < 4; ++t) {
        if (t < m.count) {
            d.pm_p[4 * k + t] = st3(m.p[t].pos, m.p[t].pen);
            d.pm_q[4 * k + t] = make_float4(m.p[t].sep, __uint_as_float(m.p[t].fid), 0.0f, 0.0f);
        } else {
            d.pm_p[4 * k + t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            d.pm_q[4 * k + t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
}
// the pair filter of world.h:1545-1577: both alive, one dynamic, layers, not switched off
__device__ __forceinline__ bool pair_collides(const DevState& d, uint4 ii, uint4 ij) {
    const uint32_t fi = d.b_flags[ii.x], fj = d.b_flags[ij.x];
    return (fi & BF_ALIVE) && (fj & BF_ALIVE) && ((fi & BF_TYPE) == 2u || (fj & BF_TYPE) == 2u) &&
           (d.layer_mask[ii.z & 31u] & (1u << (ij.z & 31u))) && !contacts_off(d, ii.x, ij.x);
}
__device__ __forceinline__ bool mesh_pair(uint4 ii, uint4 ij) {
    return int(ii.y & 15u) == SH_MESH || int(ij.y & 15u) == SH_MESH;
}

__global__ void k_narrowphase(DevState d, StepParams prm, uint32_t n_pairs) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_pairs) return;
    uint2 pr = d.pairs[k];
    uint32_t i = pr.x, j = pr.y;
    uint4 ii = d.s_info[i], ij = d.s_info[j];
    if (mesh_pair(ii, ij)) return;  // k_narrowphase_mesh's
    uint32_t bi = ii.x, bj = ij.x;
    uint32_t valid = 0;
    Man m;
    m.count = 0;
    if (pair_collides(d, ii, ij)) {
        ShapeGeom si = load_shape(d, i, ii), sj = load_shape(d, j, ij);
        float4 vi = d.b_vel[bi], vj = d.b_vel[bj];
        bool spec = (len(ld3(vi)) * prm.h > shape_min_extent(d, si)) || (len(ld3(vj)) * prm.h > shape_min_extent(d, sj));
        if (si.type == SH_HULL || sj.type == SH_HULL) {
            if (hull_contacts(hull_tab(d), si, ld3(d.b_pos[bi]), ldq(d.b_rot[bi]), sj, ld3(d.b_pos[bj]), ldq(d.b_rot[bj]), m))
                valid = 1u;
        } else if (generate_contacts(si, ld3(d.b_pos[bi]), ldq(d.b_rot[bi]), sj, ld3(d.b_pos[bj]), ldq(d.b_rot[bj]), spec, m)) {
            valid = 1u;
        }
    }
    d.pm_valid[k] = valid;
    if (valid) store_pair_manifold(d, k, i, j, ii, ij, m);
}

// The pairs with a mesh (narrowphase.h:313-320): generate_mesh_contacts with the normal turned
// from a to b; no mesh against mesh, no speculative contacts. Launched only when meshes exist.
__global__ void k_narrowphase_mesh(DevState d, uint32_t n_pairs) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_pairs) return;
    uint2 pr = d.pairs[k];
    uint32_t i = pr.x, j = pr.y;
    uint4 ii = d.s_info[i], ij = d.s_info[j];
    if (!mesh_pair(ii, ij)) return;
    uint32_t valid = 0;
    Man m;
    m.count = 0;
    if (pair_collides(d, ii, ij)) {
        const ShapeGeom si = load_shape(d, i, ii), sj = load_shape(d, j, ij);
        const v3 pi = ld3(d.b_pos[ii.x]), pj = ld3(d.b_pos[ij.x]);
        const quat ri = ldq(d.b_rot[ii.x]), rj = ldq(d.b_rot[ij.x]);
        bool overflow = false;
        if (si.type == SH_MESH && sj.type == SH_MESH) {
            valid = 0u;
        } else if (sj.type == SH_MESH) {
            if (mesh_contacts(mesh_tab(d), hull_tab(d), si, pi, ri, sj, pj, rj, m, overflow)) {
                m.n = -m.n;
                valid = 1u;
            }
        } else if (mesh_contacts(mesh_tab(d), hull_tab(d), sj, pj, rj, si, pi, ri, m, overflow)) {
            valid = 1u;
        }
        if (overflow) atomicOr(&d.cnt->err, ERR_MESH_VERTICES);
    }
    d.pm_valid[k] = valid;
    if (valid) store_pair_manifold(d, k, i, j, ii, ij, m);
}

// compaction in pair order (world.h:1647-1653)
__global__ void k_compact(DevState d, uint32_t n_pairs) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t pts = 0;
    if (k < n_pairs) {
        if (k == n_pairs - 1) {
            // clamp to the buffer so later kernels stay in bounds; the overflow is reported
            uint32_t total = d.pm_off[k] + d.pm_valid[k];
            d.cnt->manifolds = total < d.cap_manifolds ? total : d.cap_manifolds;
        }
        uint32_t m = d.pm_off[k];
        if (d.pm_valid[k] && m >= d.cap_manifolds) {
            atomicOr(&d.cnt->err, ERR_MANIFOLDS);
        } else if (d.pm_valid[k]) {
            d.m_ab[m] = d.pm_ab[k];
            d.m_shapes[m] = d.pairs[k];
            d.m_nf[m] = d.pm_nf[k];
            uint4 meta = d.pm_meta[k];
            d.m_meta[m] = meta;
            for (int t = 0; t < 4; ++t) {
                d.m_p[4 * m + t] = d.pm_p[4 * k + t];
                d.m_q[4 * m + t] = d.pm_q[4 * k + t];
            }
            pts = meta.x & 255u;
        }
    }
    pts = wave_sum(pts);
    if (lane0() && pts) atomicAdd(&d.cnt->points, pts);
}

// ------------------------------------------------------------------------------------------
// colouring: Jones-Plassmann rounds with hashed priorities, first-fit colour from per-body
// 64-bit masks. Only awake dynamic bodies constrain (apply_impulse_pair writes only those).
// ------------------------------------------------------------------------------------------
constexpr uint32_t kUncoloured = 0xFFFFFFFFu, kSkipped = 0xFFFFFFFEu, kTrigger = 0xFFFFFFFDu;

__global__ void k_colour_init(DevState d, uint32_t n_bound) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n_man = d.cnt->manifolds;
    uint32_t add = 0;
    if (m < n_man && m < n_bound) {
        uint2 ab = d.m_ab[m];
        uint4 meta = d.m_meta[m];
        bool trig = (meta.x >> 8) & 1u;
        bool act = body_active(d.b_flags[ab.x]) || body_active(d.b_flags[ab.y]);
        uint32_t c = trig ? kTrigger : (act ? kUncoloured : kSkipped);
        d.m_colour[m] = c;
        add = c == kUncoloured ? 1u : 0u;
    }
    add = wave_sum(add);
    if (lane0() && add) atomicAdd(&d.cnt->remaining, add);
}

// Jones-Plassmann priority of manifold m in a round: manifolds with more contact points first
// (they take the low colours, so the later colours hold only one-point manifolds, whose three rows
// are a quarter of a four-point manifold's twelve: a colour lasts as long as its longest
// manifold), then a hash of the pair, then the index.
__device__ __forceinline__ unsigned long long colour_key(uint32_t round, uint32_t m, uint2 ab, uint2 sh, uint32_t cnt) {
    uint32_t pr = hash32(ab.x * 0x9E3779B1u ^ hash32(ab.y + 0x85EBCA77u) ^ (sh.x * 0xC2B2AE3Du)) & 0xFFFu;
    const unsigned long long np = cnt < 4u ? cnt : 4u;
    return (static_cast<unsigned long long>(round) << 47) | (np << 44) | (static_cast<unsigned long long>(pr) << 32) | m;
}

__global__ void k_colour_propose(DevState d, uint32_t n_bound, uint32_t round) {
    if (d.cnt->remaining == 0u) return;
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_bound || m >= d.cnt->manifolds || d.m_colour[m] != kUncoloured) return;
    uint2 ab = d.m_ab[m];
    unsigned long long key = colour_key(round, m, ab, d.m_shapes[m], d.m_meta[m].x & 255u);
    if (body_active(d.b_flags[ab.x])) atomicMax(&d.col_best[ab.x], key);
    if (body_active(d.b_flags[ab.y])) atomicMax(&d.col_best[ab.y], key);
}

__global__ void k_colour_select(DevState d, uint32_t n_bound, uint32_t round) {
    if (d.cnt->remaining == 0u) return;
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t took = 0, top = 0;
    if (m < n_bound && m < d.cnt->manifolds && d.m_colour[m] == kUncoloured) {
        uint2 ab = d.m_ab[m];
        unsigned long long key = colour_key(round, m, ab, d.m_shapes[m], d.m_meta[m].x & 255u);
        bool aa = body_active(d.b_flags[ab.x]), ba = body_active(d.b_flags[ab.y]);
        bool mine = (!aa || d.col_best[ab.x] == key) && (!ba || d.col_best[ab.y] == key);
        if (mine) {
            unsigned long long used = (aa ? d.col_mask[ab.x] : 0ull) | (ba ? d.col_mask[ab.y] : 0ull);
            unsigned long long freeb = ~used;
            if (freeb == 0ull) {
                atomicOr(&d.cnt->err, ERR_COLOURS);
            } else {
                uint32_t c = uint32_t(__ffsll(static_cast<long long>(freeb)) - 1);
                d.m_colour[m] = c;
                unsigned long long bit = 1ull << c;
                if (aa) atomicOr(&d.col_mask[ab.x], bit);
                if (ba) atomicOr(&d.col_mask[ab.y], bit);
                took = 1u;
                top = c + 1u;
            }
        }
    }
    took = wave_sum(took);
    top = wave_max(top);
    if (lane0() && took) {
        atomicSub(&d.cnt->remaining, took);
        atomicMax(&d.cnt->colours, top);
    }
}

// Stable counting sort of manifolds by bucket: colour c's multi-point manifolds (2c), then its
// one-point ones (2c + 1), then the group the solver skips. The manifolds of one colour write
// disjoint bodies, so their order within the colour does not change the result; grouping them by
// point count keeps a wave's lanes on the same number of rows (a wave runs as many point rows as
// its longest manifold has, and most manifolds have one point).
__device__ __forceinline__ int bucket_of(uint32_t c, uint32_t cnt) {
    if (c < uint32_t(kMaxColours)) return int(2u * c + (cnt <= 1u ? 1u : 0u));
    return c == kSkipped ? 2 * kMaxColours : -1;
}

__global__ void k_colour_hist(DevState d, uint32_t n_man, uint32_t nblk) {
    __shared__ uint32_t hist[kSortBuckets];
    for (int b = threadIdx.x; b < kSortBuckets; b += blockDim.x) hist[b] = 0u;
    __syncthreads();
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n_man && m < d.cnt->manifolds) {
        int b = bucket_of(d.m_colour[m], d.m_meta[m].x & 255u);
        if (b >= 0) atomicAdd(&hist[b], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kSortBuckets; b += blockDim.x) d.col_table[b * nblk + blockIdx.x] = hist[b];
}
__global__ void k_colour_scatter(DevState d, uint32_t n_man, uint32_t nblk) {
    __shared__ int bk[kBlock];
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    int b = (m < n_man && m < d.cnt->manifolds) ? bucket_of(d.m_colour[m], d.m_meta[m].x & 255u) : -1;
    bk[threadIdx.x] = b;
    __syncthreads();
    if (b < 0) return;
    uint32_t rank = 0;
    for (uint32_t t = 0; t < threadIdx.x; ++t) rank += (bk[t] == b) ? 1u : 0u;
    d.order[d.col_table_off[b * nblk + blockIdx.x] + rank] = m;
}
// colour_off[c]: first solve position of colour c (c = kMaxColours: the skipped group; one past
// it: the total)
__global__ void k_colour_offsets(DevState d, uint32_t nblk, const uint32_t* total) {
    for (int c = threadIdx.x; c <= kBuckets; c += blockDim.x)
        d.cnt->colour_off[c] = c < kBuckets ? d.col_table_off[size_t(2 * c) * nblk] : *total;
    if (threadIdx.x == 0) {
        d.cnt->active = d.col_table_off[size_t(2 * kMaxColours) * nblk];
        d.cnt->ordered = *total;
    }
}

// ------------------------------------------------------------------------------------------
// solver (dynamics/solver.h)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ bool cache_lookup(const DevState& d, int table, unsigned long long key, float4& val) {
    const unsigned long long* keys = d.ck_keys[table];
    uint32_t mask = d.cache_mask;
    uint32_t idx = hash64(key) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        unsigned long long k = keys[idx];
        if (k == key) { val = d.ck_val[table][idx]; return true; }
        if (k == kEmptyKey) return false;
        idx = (idx + 1u) & mask;
    }
    return false;
}

// Gathers manifolds into solve order, captures restitution targets (solver.h:121-137) and
// seeds impulses from the previous step's cache (solver.h:144-175); misses start at zero, or, with
// the learned warm start on, queue their features for the network.
__device__ __forceinline__ FeatBody feat_body(const DevState& d, uint32_t i) {
    FeatBody b;
    uint32_t f = d.b_flags[i];
    bool act = body_active(f);
    b.x = ld3(d.b_pos[i]);
    float4 v = d.b_vel[i];
    b.v = ld3(v);
    b.w = ld3(d.b_angv[i]);
    b.im = act ? v.w : 0.0f;
    if (act) b.I = load_iiw(d.b_iiw, i);
    else { b.I.c0 = splat(0.0f); b.I.c1 = splat(0.0f); b.I.c2 = splat(0.0f); }
    return b;
}

struct SBody {
    v3 v, w, x;
    float im;
    bool act;
    m3 I;
};
__device__ __forceinline__ SBody load_sbody(const DevState& d, uint32_t i) {
    SBody b;
    float4 v = ld_dev(d.sb_v, i), w = ld_dev(d.sb_w, i);
    float4 a = d.sb_I[3 * i], c = d.sb_I[3 * i + 1], e = d.sb_I[3 * i + 2];
    b.v = ld3(v); b.im = v.w;
    b.w = ld3(w); b.act = w.w != 0.0f;
    b.I.c0 = mk(a.x, a.y, a.z); b.I.c1 = mk(a.w, c.x, c.y); b.I.c2 = mk(c.z, c.w, e.x);
    b.x = mk(e.y, e.z, e.w);
    return b;
}
__device__ __forceinline__ void store_sbody(const DevState& d, uint32_t i, const SBody& b) {
    if (!b.act) return;
    d.sb_v[i] = st3(b.v, b.im);
    d.sb_w[i] = st3(b.w, 1.0f);
}
// The per-step constants of a contact point's rows: normal and tangent effective masses
// (solver.h:226-277, the same expressions the rows evaluate) and the separation bias -sep / h used
// when the separation is positive.
__device__ __forceinline__ float4 point_row(const SBody& A, const SBody& B, v3 n, v3 pos, float sep, float h) {
    v3 ra = pos - A.x, rb = pos - B.x;
    v3 t1, t2;
    orthonormal_basis(n, t1, t2);
    v3 ran = cross(ra, n), rbn = cross(rb, n);
    float kn = A.im + B.im + dot(ran, mul(A.I, ran)) + dot(rbn, mul(B.I, rbn));
    v3 ra1 = cross(ra, t1), rb1 = cross(rb, t1);
    float k1 = A.im + B.im + dot(ra1, mul(A.I, ra1)) + dot(rb1, mul(B.I, rb1));
    v3 ra2 = cross(ra, t2), rb2 = cross(rb, t2);
    float k2 = A.im + B.im + dot(ra2, mul(A.I, ra2)) + dot(rb2, mul(B.I, rb2));
    return make_float4(kn, k1, k2, sep > 0.0f ? -sep / h : 0.0f);
}
// The normal row's velocity target (solver.h:236-240): the separation bias when the point is
// separated, else the restitution target in the relax passes and zero in the others.
__device__ __forceinline__ float4 row_for_pass(float4 row, float4 aux, int use_rest) {
    row.w = aux.x > 0.0f ? row.w : (use_rest ? aux.y : 0.0f);
    return row;
}

__global__ void k_solver_setup(DevState d, StepParams prm, uint32_t bound) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= bound || s >= d.cnt->ordered) return;
    uint32_t m = d.order[s];
    uint2 ab = d.m_ab[m];
    float4 nf = d.m_nf[m];
    uint4 meta = d.m_meta[m];
    uint32_t cnt = meta.x & 255u;
    float rest = __uint_as_float(meta.y);
    bool active = s < d.cnt->active;
    d.sm_ab[s] = ab;
    d.sm_nf[s] = nf;
    d.sm_cnt[s] = cnt;
    v3 n = ld3(nf);
    v3 pa = ld3(d.b_pos[ab.x]), pb = ld3(d.b_pos[ab.y]);
    v3 va = ld3(d.b_vel[ab.x]), vb = ld3(d.b_vel[ab.y]);
    v3 wa = ld3(d.b_angv[ab.x]), wb = ld3(d.b_angv[ab.y]);
    int prev = d.cur ^ 1;
    bool fb_loaded = false;
    FeatBody FA, FB;
    const SBody SA = load_sbody(d, ab.x), SB = load_sbody(d, ab.y);
    for (uint32_t t = 0; t < 4; ++t) {
        float4 pp = d.m_p[4 * m + t], qq = d.m_q[4 * m + t];
        float4 aux = make_float4(qq.x, 0.0f, qq.y, 0.0f);
        float4 imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (t < cnt) {
            v3 pos = ld3(pp);
            v3 ra = pos - pa, rb = pos - pb;
            v3 vpa = va + cross(wa, ra), vpb = vb + cross(wb, rb);
            float vap = dot(vpb - vpa, n);
            aux.y = (-vap > prm.rest_thresh) ? -rest * vap : 0.0f;
            if (active) {
                float4 cv;
                bool hit = cache_lookup(d, prev, pack_key(ab.x, ab.y, __float_as_uint(qq.y)), cv);
                if (hit) imp = make_float4(cv.x, cv.y, cv.z, 0.0f);
                else aux.w = 1.0f;  // warm-start miss
                if ((!hit && d.mlp) || d.record) {
                    if (!fb_loaded) { FA = feat_body(d, ab.x); FB = feat_body(d, ab.y); fb_loaded = true; }
                    float f[kMlpIn];
                    contact_features(FA, FB, n, pos, pp.w, qq.x, nf.w, rest, f);
                    float* row;
                    if (d.record) {
                        row = d.feat + size_t(4 * s + t) * kMlpIn;
                    } else {
                        uint32_t k = atomicAdd(&d.cnt->predicted, 1u);
                        d.miss_idx[k] = 4 * s + t;
                        row = d.feat + size_t(k) * kMlpIn;
                    }
                    for (int q = 0; q < kMlpIn; ++q) row[q] = f[q];
                }
            }
        }
        d.sp_pos[4 * s + t] = pp;
        d.sp_aux[4 * s + t] = aux;
        d.sp_imp[4 * s + t] = imp;
        d.sp_row[4 * s + t] = t < cnt ? point_row(SA, SB, n, ld3(pp), aux.x, prm.h) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// First colour of the solve's tail: the trailing colours of at most one wave each (at least two
// of them), which one wave runs back to back; C when there is no such tail.
__device__ __forceinline__ uint32_t tail_start(const uint32_t* co, uint32_t C) {
    uint32_t t = C;
    while (t > 0 && co[t] - co[t - 1] <= 64u) --t;
    return C - t >= 2u ? t : C;
}
// Residual mode: a manifold's end-of-pass residual reads only its bodies' velocities. When no tail
// colour writes either (awake dynamic) body, they are final before the tail starts, and the
// residual is computed by the waves the tail leaves idle.
__device__ __forceinline__ bool sb_active(const DevState& d, uint32_t i) { return d.sb_w[i].w != 0.0f; }
__global__ void k_ready_last(DevState d, uint32_t bound) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= bound || s >= d.cnt->active) return;
    const uint32_t c = d.m_colour[d.order[s]];
    const uint2 ab = d.sm_ab[s];
    if (sb_active(d, ab.x)) atomicMax(&d.body_last[ab.x], c);
    if (sb_active(d, ab.y)) atomicMax(&d.body_last[ab.y], c);
}
// the order within each part depends on timing; nothing depends on it (the residual is a maximum)
__global__ void k_ready_split(DevState d, uint32_t bound) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = d.cnt->active;
    if (s >= bound || s >= n) return;
    const uint32_t T = tail_start(d.cnt->colour_off, d.cnt->colours);
    const uint2 ab = d.sm_ab[s];
    const bool late = (sb_active(d, ab.x) && d.body_last[ab.x] >= T) || (sb_active(d, ab.y) && d.body_last[ab.y] >= T);
    if (late) d.ready_list[n - 1u - atomicAdd(&d.ready_cnt[1], 1u)] = s;
    else d.ready_list[atomicAdd(&d.ready_cnt[0], 1u)] = s;
}

// clamps each prediction to the friction cone (lambda_n >= 0, |lambda_t| <= mu lambda_n) and makes
// it the point's initial impulse
__global__ void k_apply_predictions(DevState d, uint32_t n_bound) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_bound || j >= d.cnt->predicted) return;
    uint32_t p = d.miss_idx[j];
    float mu = d.sm_nf[p / 4].w;
    float ln = fmaxf(d.pred[3 * j], 0.0f);
    float lim = mu * ln;
    float t1 = fminf(fmaxf(d.pred[3 * j + 1], -lim), lim);
    float t2 = fminf(fmaxf(d.pred[3 * j + 2], -lim), lim);
    d.pred[3 * j] = ln;
    d.pred[3 * j + 1] = t1;
    d.pred[3 * j + 2] = t2;
    d.sp_imp[p] = make_float4(ln, t1, t2, 0.0f);
}

// compact solver bodies: effective inverse mass / inertia are zero unless dynamic and awake
__global__ void k_solver_bodies(DevState d) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nb) return;
    uint32_t f = d.b_flags[i];
    bool act = body_active(f);
    float4 v = d.b_vel[i], w = d.b_angv[i], p = d.b_pos[i];
    d.sb_v[i] = make_float4(v.x, v.y, v.z, act ? v.w : 0.0f);
    d.sb_w[i] = make_float4(w.x, w.y, w.z, act ? 1.0f : 0.0f);
    m3 I;
    if (act) I = load_iiw(d.b_iiw, i);
    else { I.c0 = splat(0.0f); I.c1 = splat(0.0f); I.c2 = splat(0.0f); }
    store_iiw(d.sb_I, i, I, p.x, p.y, p.z);
}

// apply_impulse_pair (solver.h:86-95)
__device__ __forceinline__ void apply_pair(SBody& a, SBody& b, v3 p, v3 J) {
    if (a.act) { a.v = a.v - J * a.im; a.w = a.w - mul(a.I, cross(p - a.x, J)); }
    if (b.act) { b.v = b.v + J * b.im; b.w = b.w + mul(b.I, cross(p - b.x, J)); }
}

// Solver manifolds load all their data in one go: every load of a slot is independent of the
// others, so they are all in flight together (a point loop that loads as it goes waits on each
// point's loads in turn, since its stores may alias the next point's loads). Slots are padded
// to four points, so the loads need no count. The four points are separate SSA values and a
// point chosen at run time is a select (sel4), never an indexed access that would put the
// slot in scratch memory.
__device__ __forceinline__ float4 sel4(uint32_t i, float4 a, float4 b, float4 c, float4 e) {
    return i == 0u ? a : i == 1u ? b : i == 2u ? c : e;
}

// warm_start's impulse application (solver.h:164-173) for solver manifold s
__device__ __forceinline__ void warm_manifold(const DevState& d, uint32_t s) {
    const uint2 ab = d.sm_ab[s];
    const float4 nf = d.sm_nf[s];
    const uint32_t cnt = d.sm_cnt[s];
    const float4 p0 = d.sp_pos[4 * s], p1 = d.sp_pos[4 * s + 1], p2 = d.sp_pos[4 * s + 2], p3 = d.sp_pos[4 * s + 3];
    const float4 m0 = d.sp_imp[4 * s], m1 = d.sp_imp[4 * s + 1], m2 = d.sp_imp[4 * s + 2], m3v = d.sp_imp[4 * s + 3];
    SBody A = load_sbody(d, ab.x), B = load_sbody(d, ab.y);
    v3 n = ld3(nf);
    v3 t1, t2;
    orthonormal_basis(n, t1, t2);
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        if (i >= cnt) break;
        float4 imp = sel4(i, m0, m1, m2, m3v);
        v3 J = n * imp.x + t1 * imp.y + t2 * imp.z;
        if (dot(J, J) > 0.0f) apply_pair(A, B, ld3(sel4(i, p0, p1, p2, p3)), J);
    }
    store_sbody(d, ab.x, A);
    store_sbody(d, ab.y, B);
}
__global__ void k_warm_apply(DevState d, uint32_t begin, uint32_t end) {
    uint32_t s = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (s < end) warm_manifold(d, s);
}

// solve_velocity_pass (solver.h:212-281), one colour
__device__ __forceinline__ void velocity_manifold(const DevState& d, const StepParams& prm, uint32_t s, int use_rest, uint32_t offset);

__global__ void __launch_bounds__(kBlock) k_velocity(DevState d, StepParams prm, uint32_t begin, uint32_t end, int use_rest, uint32_t offset) {
    uint32_t s = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (prm.profile) {
        // algorithmic bytes of one velocity pass: 192 per manifold + 40 per point
        uint32_t bytes = s < end ? 192u + 40u * d.sm_cnt[s] : 0u;
        bytes = wave_sum(bytes);
        if (lane0() && bytes) atomicAdd(&d.cnt->prof_vel_bytes, static_cast<unsigned long long>(bytes));
    }
    if (s >= end) return;
    velocity_manifold(d, prm, s, use_rest, offset);
}

// One contact point's three rows (solver.h:226-277). Returns the updated impulses. The effective
// masses depend only on the bodies' inverse mass, inverse inertia and position and on the point,
// none of which a velocity pass changes, so they are computed once per step (point_row, with the
// same expressions the rows would evaluate); row = (kn, kt1, kt2, the normal row's target).
__device__ __forceinline__ float4 velocity_point(SBody& A, SBody& B, v3 n, v3 t1, v3 t2, float mu, v3 pos,
                                                 float4 row, float4 imp) {
    v3 ra = pos - A.x, rb = pos - B.x;
    {
        const float kn = row.x;
        if (kn > kEps) {
            v3 va = A.v + cross(A.w, ra), vb = B.v + cross(B.w, rb);
            float vn = dot(vb - va, n);
            float target = row.w;
            float lambda = -(vn - target) / kn;
            float ni = gmax(imp.x + lambda, 0.0f);
            float delta = ni - imp.x;
            imp.x = ni;
            apply_pair(A, B, pos, n * delta);
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        v3 dir = t == 0 ? t1 : t2;
        const float kt = t == 0 ? row.y : row.z;
        if (kt <= kEps) continue;
        v3 va = A.v + cross(A.w, ra), vb = B.v + cross(B.w, rb);
        float vt = dot(vb - va, dir);
        float lt = -vt / kt;
        float maxf = mu * imp.x;
        float old = t == 0 ? imp.y : imp.z;
        float nt = sclamp(old + lt, -maxf, maxf);
        float delta = nt - old;
        if (t == 0) imp.y = nt; else imp.z = nt;
        apply_pair(A, B, pos, dir * delta);
    }
    return imp;
}

// Points run in the order (k + offset) % count (solver.h:221-224). The k-th point's slot index
// is computed first, so the loads of all points are issued together from global memory and the
// loop itself works on fixed registers.
__device__ __forceinline__ void velocity_manifold(const DevState& d, const StepParams& prm, uint32_t s, int use_rest, uint32_t offset) {
    const uint2 ab = d.sm_ab[s];
    const float4 nf = d.sm_nf[s];
    const uint32_t cnt = d.sm_cnt[s];
    const uint32_t r = cnt ? offset % cnt : 0u;
    uint32_t idx[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) idx[k] = 4u * s + (k < cnt ? (k + r) % cnt : k);
    float4 pp[4], row[4], imp[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        pp[k] = row[k] = imp[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (k < cnt) {
            pp[k] = d.sp_pos[idx[k]];
            row[k] = row_for_pass(d.sp_row[idx[k]], d.sp_aux[idx[k]], use_rest);
            imp[k] = ld_dev(d.sp_imp, idx[k]);
        }
    }
    SBody A = load_sbody(d, ab.x), B = load_sbody(d, ab.y);
    v3 n = ld3(nf);
    v3 t1, t2;
    orthonormal_basis(n, t1, t2);
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k >= cnt) break;
        imp[k] = velocity_point(A, B, n, t1, t2, nf.w, ld3(pp[k]), row[k], imp[k]);
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
        if (k < cnt) d.sp_imp[idx[k]] = imp[k];
    store_sbody(d, ab.x, A);
    store_sbody(d, ab.y, B);
}

// Velocity residual (m/s) of solver manifold s: normal rows measure the complementarity
// violation of (lambda_n, v_n - target); friction rows the violation of stick (|lambda_t| <
// mu lambda_n needs v_t = 0) or slip (the friction impulse must oppose v_t).
__device__ __forceinline__ float residual_manifold(const DevState& d, const StepParams& prm, uint32_t s, int use_rest) {
    float r = 0.0f;
    {
        uint2 ab = d.sm_ab[s];
        SBody A = load_sbody(d, ab.x), B = load_sbody(d, ab.y);
        float4 nf = d.sm_nf[s];
        v3 n = ld3(nf);
        v3 t1, t2;
        orthonormal_basis(n, t1, t2);
        uint32_t cnt = d.sm_cnt[s];
        // all the points' loads go out together; a wave whose manifolds all have fewer points
        // skips the loads of the points none of them has
        float4 pp4[4], aux4[4], imp4[4];
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            pp4[i] = aux4[i] = imp4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (i < cnt) {
                pp4[i] = d.sp_pos[4 * s + i];
                aux4[i] = d.sp_aux[4 * s + i];
                imp4[i] = ld_dev(d.sp_imp, 4 * s + i);
            }
        }
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            if (i >= cnt) break;
            const float4 pp = pp4[i], aux = aux4[i], imp = imp4[i];
            v3 pos = ld3(pp);
            v3 ra = pos - A.x, rb = pos - B.x;
            v3 dv = (B.v + cross(B.w, rb)) - (A.v + cross(A.w, ra));
            float vn = dot(dv, n);
            float target = aux.x > 0.0f ? -aux.x / prm.h : (use_rest ? aux.y : 0.0f);
            float rn = imp.x > 0.0f ? fabsf(vn - target) : gmax(0.0f, target - vn);
            r = gmax(r, rn);
            float maxf = nf.w * imp.x;
            for (int t = 0; t < 2; ++t) {
                float vt = dot(dv, t == 0 ? t1 : t2);
                float lt = t == 0 ? imp.y : imp.z;
                float rt = 0.0f;
                if (maxf > 0.0f) rt = fabsf(lt) < maxf ? fabsf(vt) : gmax(0.0f, lt < 0.0f ? -vt : vt);
                r = gmax(r, rt);
            }
        }
    }
    return r;
}
// max-reduces the residual of the active solver manifolds into cnt->residual_bits
__global__ void k_residual(DevState d, StepParams prm, uint32_t n_active, int use_rest) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    float r = s < n_active ? residual_manifold(d, prm, s, use_rest) : 0.0f;
    for (int o = 32; o > 0; o >>= 1) r = fmaxf(r, __shfl_xor(r, o, 64));
    if ((threadIdx.x & 63) == 0 && r > 0.0f) atomicMax(&d.cnt->residual_bits, __float_as_uint(r));
}
__global__ void k_residual_reset(DevState d) { d.cnt->residual_bits = 0u; }
__global__ void k_set_passes(DevState d, uint32_t n) { d.cnt->passes = n; }

// solve_position setup: positions and reference snapshot (solver.h:295-296)
__global__ void k_position_setup(DevState d) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nb) return;
    float4 p = d.b_pos[i];
    d.sb_x[i] = make_float4(p.x, p.y, p.z, d.sb_v[i].w);
    d.sb_ref[i] = p;
}
// solve_position for solver manifold s (solver.h:298-327)
__device__ __forceinline__ void position_manifold(const DevState& d, const StepParams& prm, uint32_t s) {
    uint2 ab = d.sm_ab[s];
    v3 n = ld3(d.sm_nf[s]);
    uint32_t cnt = d.sm_cnt[s];
    float pen4[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) pen4[t] = d.sp_pos[4 * s + t].w;
    float4 xa4 = ld_dev(d.sb_x, ab.x), xb4 = ld_dev(d.sb_x, ab.y);
    bool aa = ld_dev(d.sb_w, ab.x).w != 0.0f, ba = ld_dev(d.sb_w, ab.y).w != 0.0f;
    v3 ra = ld3(ld_dev(d.sb_ref, ab.x)), rb = ld3(ld_dev(d.sb_ref, ab.y));
    float ima = xa4.w, imb = xb4.w;
    float k = ima + imb;
    if (k <= kEps) return;
    v3 xa = ld3(xa4), xb = ld3(xb4);
    v3 da = xa - ra, db = xb - rb;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        if (i >= cnt) break;
        float pen = pen4[i];
        float cur = pen - dot(db - da, n);
        float mag = gmax(0.0f, cur - prm.slop) * prm.pos_corr;
        mag = gmin(mag, prm.max_corr);
        if (mag <= 0.0f) continue;
        v3 corr = n * (mag / k);
        if (aa) { xa = xa - corr * ima; da = xa - ra; }
        if (ba) { xb = xb + corr * imb; db = xb - rb; }
    }
    if (aa) d.sb_x[ab.x] = st3(xa, ima);
    if (ba) d.sb_x[ab.y] = st3(xb, imb);
}
__global__ void k_position(DevState d, StepParams prm, uint32_t begin, uint32_t end) {
    uint32_t s = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (s < end) position_manifold(d, prm, s);
}

// write solver bodies back; rebuild the warm-start cache (solver.h:767-781): insert in solve
// order, last writer (highest order) wins on key collisions
__global__ void k_bodies_back(DevState d) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nb) return;
    if (d.sb_w[i].w == 0.0f) return;
    float4 v = d.sb_v[i], w = d.sb_w[i], x = d.sb_x[i];
    float4 ov = d.b_vel[i], ow = d.b_angv[i], op = d.b_pos[i];
    d.b_vel[i] = make_float4(v.x, v.y, v.z, ov.w);
    d.b_angv[i] = make_float4(w.x, w.y, w.z, ow.w);
    d.b_pos[i] = make_float4(x.x, x.y, x.z, op.w);
}
// position of point i of the manifold of shapes sh in the reference's insertion order (pair
// order, then point index), plus one; 0 marks an empty slot
__device__ __forceinline__ unsigned long long ref_order(uint2 sh, uint32_t i) {
    return ((static_cast<unsigned long long>(sh.x) << 40) | (static_cast<unsigned long long>(sh.y) << 4) | i) + 1ull;
}
__global__ void k_cache_insert(DevState d, uint32_t bound) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= bound || s >= d.cnt->ordered) return;
    uint2 ab = d.sm_ab[s];
    uint32_t cnt = d.sm_cnt[s];
    unsigned long long* keys = d.ck_keys[d.cur];
    unsigned long long* ord = d.ck_ord[d.cur];
    const uint2 sh = d.m_shapes[d.order[s]];
    uint32_t mask = d.cache_mask;
    for (uint32_t i = 0; i < cnt; ++i) {
        unsigned long long key = pack_key(ab.x, ab.y, __float_as_uint(d.sp_aux[4 * s + i].z));
        uint32_t idx = hash64(key) & mask;
        bool done = false;
        for (uint32_t probe = 0; probe <= mask; ++probe) {
            unsigned long long old = atomicCAS(&keys[idx], kEmptyKey, key);
            if (old == kEmptyKey || old == key) {
                atomicMax(&ord[idx], ref_order(sh, i));
                done = true;
                break;
            }
            idx = (idx + 1u) & mask;
        }
        if (!done) atomicOr(&d.cnt->err, ERR_CACHE);
    }
}
// the value's w keeps the manifold's colour: the next step's colouring starts from it
__global__ void k_cache_write(DevState d, uint32_t bound) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= bound || s >= d.cnt->ordered) return;
    const float colour = __uint_as_float(d.m_colour[d.order[s]]);
    uint2 ab = d.sm_ab[s];
    uint32_t cnt = d.sm_cnt[s];
    const unsigned long long* keys = d.ck_keys[d.cur];
    const unsigned long long* ord = d.ck_ord[d.cur];
    const uint2 sh = d.m_shapes[d.order[s]];
    uint32_t mask = d.cache_mask;
    for (uint32_t i = 0; i < cnt; ++i) {
        unsigned long long key = pack_key(ab.x, ab.y, __float_as_uint(d.sp_aux[4 * s + i].z));
        uint32_t idx = hash64(key) & mask;
        for (uint32_t probe = 0; probe <= mask; ++probe) {
            if (keys[idx] == key) {
                if (ord[idx] == ref_order(sh, i)) {
                    float4 v = d.sp_imp[4 * s + i];
                    d.ck_val[d.cur][idx] = make_float4(v.x, v.y, v.z, colour);
                }
                break;
            }
            if (keys[idx] == kEmptyKey) break;
            idx = (idx + 1u) & mask;
        }
    }
}

// ------------------------------------------------------------------------------------------
// islands and sleep (solver.h:808-876, island.h:25-49)
// ------------------------------------------------------------------------------------------
// Parent pointers only ever move to smaller indices (a root is hooked under the smaller root),
// so a stale read is an older ancestor: following it still ends in the same set. Reads are plain
// loads; the only atomic is the hook, whose returned value continues the search when another
// thread hooked the root first.
__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
    while (true) {
        uint32_t p = parent[x];
        if (p == x) return x;
        uint32_t g = parent[p];
        if (g != p) parent[x] = g;  // path halving (a benign race: g is an ancestor of x)
        x = p;
    }
}
__global__ void k_uf_init(DevState d) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nb) return;
    d.uf_parent[i] = i;
    d.isl_min[i] = 0x7F800000u;
}
// unite dynamic-dynamic manifolds (triggers included); roots hook onto the smaller index so the
// final root is the component's smallest body index
__global__ void k_uf_union(DevState d, uint32_t bound) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= bound || m >= d.cnt->manifolds) return;
    uint2 ab = d.m_ab[m];
    if ((d.b_flags[ab.x] & BF_TYPE) != 2u || (d.b_flags[ab.y] & BF_TYPE) != 2u) return;
    uint32_t a = uf_find(d.uf_parent, ab.x), b = uf_find(d.uf_parent, ab.y);
    for (int guard = 0; guard < (1 << 24) && a != b; ++guard) {
        if (a > b) { uint32_t t = a; a = b; b = t; }
        uint32_t old = atomicCAS(&d.uf_parent[b], b, a);
        if (old == b) return;
        // b was hooked meanwhile: continue from its current ancestors
        b = uf_find(d.uf_parent, old);
        a = uf_find(d.uf_parent, a);
    }
}
// Each body's root, into a separate array: the parent array is only read here. (A find that
// halves paths while other threads store their final roots could overwrite a stored root with an
// intermediate ancestor, splitting one island's key in two.)
__global__ void k_uf_roots(DevState d) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nb) return;
    uint32_t x = i;
    for (uint32_t p = d.uf_parent[x]; p != x; p = d.uf_parent[x]) x = p;
    d.uf_root[i] = x;
}
// joint pairs unite islands too (solver.h:820-825)
__global__ void k_uf_union_joints(DevState d) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.nj) return;
    uint4 m = d.j_meta[j];
    if (!m.w) return;
    if ((d.b_flags[m.y] & BF_TYPE) != 2u || (d.b_flags[m.z] & BF_TYPE) != 2u) return;
    uint32_t a = uf_find(d.uf_parent, m.y), b = uf_find(d.uf_parent, m.z);
    for (int guard = 0; guard < (1 << 24) && a != b; ++guard) {
        if (a > b) { uint32_t t = a; a = b; b = t; }
        uint32_t old = atomicCAS(&d.uf_parent[b], b, a);
        if (old == b) return;
        b = uf_find(d.uf_parent, old);
        a = uf_find(d.uf_parent, a);
    }
}
__global__ void k_sleep_timers(DevState d, StepParams p) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nb) return;
    uint32_t f = d.b_flags[i];
    if (!body_active(f)) return;
    v3 v = ld3(d.b_vel[i]), w = ld3(d.b_angv[i]);
    float4 pt = d.b_pos[i];
    float l2 = dot(v, v), a2 = dot(w, w);
    if (l2 < p.sleep_lin * p.sleep_lin && a2 < p.sleep_ang * p.sleep_ang) pt.w = pt.w + p.h;
    else pt.w = 0.0f;
    d.b_pos[i] = pt;
}
// wake_if_kinematic_moving both ways per valid manifold (solver.h:787-795, 849-853)
__global__ void k_kin_wake(DevState d, StepParams p, uint32_t bound) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= bound || m >= d.cnt->manifolds) return;
    uint2 ab = d.m_ab[m];
    for (int side = 0; side < 2; ++side) {
        uint32_t kin = side == 0 ? ab.x : ab.y, dyn = side == 0 ? ab.y : ab.x;
        uint32_t fk = d.b_flags[kin], fd = d.b_flags[dyn];
        if ((fk & BF_TYPE) != 1u || (fd & BF_TYPE) != 2u || (fd & BF_AWAKE)) continue;
        v3 kv = ld3(d.b_vel[kin]), kw = ld3(d.b_angv[kin]);
        bool lm = dot(kv, kv) > p.sleep_lin * p.sleep_lin;
        bool am = dot(kw, kw) > p.sleep_ang * p.sleep_ang;
        if (lm || am) {
            atomicOr(&d.b_flags[dyn], BF_AWAKE);
            float4 pt = d.b_pos[dyn];
            pt.w = 0.0f;
            d.b_pos[dyn] = pt;
        }
    }
}
// wake_if_kinematic_moving both ways per live joint (solver.h:854-858)
__global__ void k_kin_wake_joints(DevState d, StepParams p) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.nj) return;
    uint4 m = d.j_meta[j];
    if (!m.w) return;
    for (int side = 0; side < 2; ++side) {
        uint32_t kin = side == 0 ? m.y : m.z, dyn = side == 0 ? m.z : m.y;
        uint32_t fk = d.b_flags[kin], fd = d.b_flags[dyn];
        if ((fk & BF_TYPE) != 1u || (fd & BF_TYPE) != 2u || (fd & BF_AWAKE)) continue;
        v3 kv = ld3(d.b_vel[kin]), kw = ld3(d.b_angv[kin]);
        bool lm = dot(kv, kv) > p.sleep_lin * p.sleep_lin;
        bool am = dot(kw, kw) > p.sleep_ang * p.sleep_ang;
        if (lm || am) {
            atomicOr(&d.b_flags[dyn], BF_AWAKE);
            float4 pt = d.b_pos[dyn];
            pt.w = 0.0f;
            d.b_pos[dyn] = pt;
        }
    }
}
// per-island minimum sleep timer; lanes of a wave that share a root (one big island is common)
// combine their minimum before a single atomic
__global__ void k_island_min(DevState d) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool act = i < d.nb && body_active(d.b_flags[i]);
    uint32_t root = act ? d.uf_root[i] : 0xFFFFFFFFu;
    uint32_t val = act ? __float_as_uint(d.b_pos[i].w) : 0xFFFFFFFFu;
    unsigned long long pending = __ballot(act);
    int lane = threadIdx.x & 63;
    while (pending) {
        int leader = __ffsll(static_cast<long long>(pending)) - 1;
        uint32_t r = __shfl(root, leader, 64);
        bool mine = act && root == r;
        uint32_t m = wave_min(mine ? val : 0xFFFFFFFFu);
        if (lane == leader) atomicMin(&d.isl_min[r], m);
        pending &= ~__ballot(mine);
    }
}
__global__ void k_sleep(DevState d, StepParams p) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nb) return;
    uint32_t f = d.b_flags[i];
    if (!body_active(f)) return;
    if (__uint_as_float(d.isl_min[d.uf_root[i]]) >= p.sleep_time) {
        d.b_flags[i] = f & ~BF_AWAKE;
        float4 v = d.b_vel[i], w = d.b_angv[i];
        d.b_vel[i] = make_float4(0.0f, 0.0f, 0.0f, v.w);
        d.b_angv[i] = make_float4(0.0f, 0.0f, 0.0f, w.w);
    }
}

// prev pose, integrate_velocities, update_world_inertia (world.h:1303-1309, integrator.h:109-133)
__global__ void k_integrate_velocities(DevState d, StepParams p) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t f = i < d.nb ? d.b_flags[i] : 0u;
    uint32_t aw = ((f & BF_ALIVE) && (f & BF_TYPE) == 2u && (f & BF_AWAKE)) ? 1u : 0u;
    aw = wave_sum(aw);
    if (lane0() && aw) atomicAdd(&d.cnt->awake, aw);
    if (i >= d.nb) return;
    if (!(f & BF_ALIVE)) return;
    float4 pos = d.b_pos[i], rot = d.b_rot[i];
    d.b_prev_pos[i] = make_float4(pos.x, pos.y, pos.z, 0.0f);
    d.b_prev_rot[i] = rot;
    bool dyn = (f & BF_TYPE) == 2u;
    if (dyn && (f & BF_AWAKE)) {
        float4 vm = d.b_vel[i];
        v3 v = ld3(vm);
        float s2 = dot(v, v);
        float m2 = p.max_lin_vel * p.max_lin_vel;
        if (s2 > m2 && s2 > kEps) {
            v = v * (p.max_lin_vel / sqrtf(s2));
            d.b_vel[i] = st3(v, vm.w);
        }
        v3 x = ld3(pos) + v * p.h;
        d.b_pos[i] = st3(x, pos.w);
        quat q = ldq(rot);
        v3 w = ld3(d.b_angv[i]);
        if (p.exact_quat) {
            float angle = len(w) * p.h;
            if (angle > kEps) {
                v3 ax = w / len(w);
                float half = angle * 0.5f;
                float sn = sinf(half);
                quat dq;
                dq.w = cosf(half); dq.x = ax.x * sn; dq.y = ax.y * sn; dq.z = ax.z * sn;
                q = qnormalize(qmul(dq, q));
            }
        } else {
            quat wq;
            wq.w = 0.0f; wq.x = w.x; wq.y = w.y; wq.z = w.z;
            quat dq = qmul(wq, q);
            float hh = 0.5f * p.h;
            quat s;
            s.w = q.w + dq.w * hh; s.x = q.x + dq.x * hh; s.y = q.y + dq.y * hh; s.z = q.z + dq.z * hh;
            q = qnormalize(s);
        }
        d.b_rot[i] = stq(q);
        store_iiw(d.b_iiw, i, world_inertia(q, ld3(d.b_iil[i])));
    } else if (!dyn) {
        m3 z;
        z.c0 = splat(0.0f); z.c1 = splat(0.0f); z.c2 = splat(0.0f);
        store_iiw(d.b_iiw, i, z);
    }
}

// ------------------------------------------------------------------------------------------
// joints (dynamics/solver.h:330-761, world.h:1310-1314). Each pass visits the joints layer by
// layer; a layer holds joints that share no dynamic body, in table order, so the per-body
// update order is the reference's sequential one.
// ------------------------------------------------------------------------------------------
// bodies as the velocity passes and the warm start see them: the solver bodies
__device__ __forceinline__ JBody jb_solver(const DevState& d, uint32_t i) {
    JBody b;
    float4 v = ld_dev(d.sb_v, i), w = ld_dev(d.sb_w, i);
    float4 a = d.sb_I[3 * i], c = d.sb_I[3 * i + 1], e = d.sb_I[3 * i + 2];
    b.v = ld3(v); b.im = v.w;
    b.w = ld3(w); b.act = w.w != 0.0f;
    b.I.c0 = mk(a.x, a.y, a.z); b.I.c1 = mk(a.w, c.x, c.y); b.I.c2 = mk(c.z, c.w, e.x);
    b.x = mk(e.y, e.z, e.w);
    b.q = ldq(ld_dev(d.b_rot, i));
    return b;
}
__device__ __forceinline__ void jb_solver_store(const DevState& d, uint32_t i, const JBody& b) {
    if (!b.act) return;
    d.sb_v[i] = st3(b.v, b.im);
    d.sb_w[i] = st3(b.w, 1.0f);
}
// the position pass of the solve: positions from the position-pass state
__device__ __forceinline__ JBody jb_position(const DevState& d, uint32_t i) {
    JBody b = jb_solver(d, i);
    b.x = ld3(ld_dev(d.sb_x, i));
    return b;
}
__device__ __forceinline__ void jb_position_store(const DevState& d, uint32_t i, const JBody& b) {
    if (!b.act) return;
    d.sb_x[i] = st3(b.x, b.im);
    d.b_rot[i] = stq(b.q);
}
// after integration: the body arrays
__device__ __forceinline__ JBody jb_world(const DevState& d, uint32_t i) {
    JBody b;
    uint32_t f = d.b_flags[i];
    b.act = body_active(f);
    b.x = ld3(d.b_pos[i]);
    b.q = ldq(d.b_rot[i]);
    b.v = ld3(d.b_vel[i]);
    b.w = ld3(d.b_angv[i]);
    b.im = b.act ? d.b_vel[i].w : 0.0f;
    b.I = b.act ? load_iiw(d.b_iiw, i) : zero3();
    return b;
}
__device__ __forceinline__ void jb_world_store(const DevState& d, uint32_t i, const JBody& b) {
    if (!b.act) return;
    float4 p = d.b_pos[i];
    d.b_pos[i] = st3(b.x, p.w);
    d.b_rot[i] = stq(b.q);
}
__device__ __forceinline__ float inv_step(float h) { return h > 0.0f ? 1.0f / h : 0.0f; }

__device__ __forceinline__ void joint_warm_one(const DevState& d, uint32_t j) {
    uint4 m = d.j_meta[j];
    if (!m.w) return;
    JDef def = load_jdef(d.j_def, j);
    JAcc acc = load_jacc(d.j_acc, j);
    acc.limit = 0.0f;  // limits start every step cold (solver.h:487-488)
    acc.swing = 0.0f;
    JBody a = jb_solver(d, m.y), b = jb_solver(d, m.z);
    joint_warm(def, acc, a, b);
    store_jacc(d.j_acc, j, acc);
    jb_solver_store(d, m.y, a);
    jb_solver_store(d, m.z, b);
}
__device__ __forceinline__ void joint_velocity_one(const DevState& d, uint32_t j, float inv_h) {
    uint4 m = d.j_meta[j];
    if (!m.w) return;
    JDef def = load_jdef(d.j_def, j);
    JAcc acc = load_jacc(d.j_acc, j);
    JBody a = jb_solver(d, m.y), b = jb_solver(d, m.z);
    if (!a.act && !b.act) return;
    joint_velocity(def, acc, a, b, inv_h);
    store_jacc(d.j_acc, j, acc);
    jb_solver_store(d, m.y, a);
    jb_solver_store(d, m.z, b);
}
__device__ __forceinline__ void joint_position_one(const DevState& d, const StepParams& p, uint32_t j) {
    uint4 m = d.j_meta[j];
    if (!m.w) return;
    JDef def = load_jdef(d.j_def, j);
    JBody a = jb_position(d, m.y), b = jb_position(d, m.z);
    if (!a.act && !b.act) return;
    joint_position(def, a, b, p.slop, p.pos_corr, p.max_corr);
    jb_position_store(d, m.y, a);
    jb_position_store(d, m.z, b);
}
__device__ void joint_post_one(const DevState& d, const StepParams& p, uint32_t j) {
    uint4 m = d.j_meta[j];
    if (!m.w) return;
    JDef def = load_jdef(d.j_def, j);
    JBody a = jb_world(d, m.y), b = jb_world(d, m.z);
    if (!a.act && !b.act) return;
    joint_position(def, a, b, p.slop, p.pos_corr, p.max_corr);
    jb_world_store(d, m.y, a);
    jb_world_store(d, m.z, b);
}
// the joints of layer l (forward or reverse order), strided over `nthreads` threads
__device__ __forceinline__ void joint_layer_range(const DevState& d, uint32_t l, bool rev, uint32_t& begin, uint32_t& end,
                                                  const uint32_t*& order) {
    const uint32_t* off = d.j_layer_off + (rev ? d.nj_layers + 1u : 0u);
    order = d.j_order + (rev ? d.nj : 0u);
    begin = off[l];
    end = off[l + 1];
}

// wake_jointed_islands (solver.h:887-901): sweeps until nothing changes. One block; the fixed
// point is the same whatever order the joints are visited in (every dynamic body jointed,
// through dynamic bodies, to an awake one wakes).
__global__ void __launch_bounds__(1024) k_wake_jointed(DevState d) {
    for (uint32_t pass = 0; pass <= d.nj; ++pass) {
        int changed = 0;
        for (uint32_t j = threadIdx.x; j < d.nj; j += blockDim.x) {
            uint4 m = d.j_meta[j];
            if (!m.w) continue;
            uint32_t fa = d.b_flags[m.y], fb = d.b_flags[m.z];
            if ((fa & BF_TYPE) != 2u || (fb & BF_TYPE) != 2u) continue;
            bool aa = (fa & BF_AWAKE) != 0u, ba = (fb & BF_AWAKE) != 0u;
            if (aa == ba) continue;
            uint32_t s = aa ? m.z : m.y;
            d.b_flags[s] = (aa ? fb : fa) | BF_AWAKE;
            float4 pt = d.b_pos[s];
            pt.w = 0.0f;
            d.b_pos[s] = pt;
            changed = 1;
        }
        if (!__syncthreads_or(changed)) break;
    }
}
__global__ void k_joint_warm(DevState d, uint32_t begin, uint32_t end) {
    uint32_t t = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (t < end) joint_warm_one(d, d.j_order[t]);
}
__global__ void k_joint_velocity(DevState d, StepParams p, int rev, uint32_t begin, uint32_t end) {
    uint32_t t = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (t < end) joint_velocity_one(d, d.j_order[(rev ? d.nj : 0u) + t], inv_step(p.h));
}
__global__ void k_joint_position(DevState d, StepParams p, uint32_t begin, uint32_t end) {
    uint32_t t = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (t < end) joint_position_one(d, p, d.j_order[t]);
}
// solve_joints_post_integrate (solver.h:755-761): max(1, position_iterations) forward passes on
// the integrated poses, one block, layers separated by block barriers
__global__ void __launch_bounds__(1024) k_joints_post(DevState d, StepParams p, uint32_t iters) {
    for (uint32_t it = 0; it < iters; ++it) {
        for (uint32_t l = 0; l < d.nj_layers; ++l) {
            const uint32_t b = d.j_layer_off[l], e = d.j_layer_off[l + 1];
            for (uint32_t t = b + threadIdx.x; t < e; t += blockDim.x) joint_post_one(d, p, d.j_order[t]);
            __syncthreads();
        }
    }
}
// update_world_inertia for every alive body (world.h:1313)
__global__ void k_update_inertia(DevState d) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nb) return;
    uint32_t f = d.b_flags[i];
    if (!(f & BF_ALIVE)) return;
    if ((f & BF_TYPE) == 2u) store_iiw(d.b_iiw, i, world_inertia(ldq(d.b_rot[i]), ld3(d.b_iil[i])));
    else store_iiw(d.b_iiw, i, zero3());
}

// ------------------------------------------------------------------------------------------
// persistent kernels: one cooperative launch per solve (and per colouring), grid-wide barriers
// between colours. Every wait is bounded: a barrier that does not complete within the spin
// budget raises ERR_BARRIER and releases all waiters, so the kernel always terminates.
// ------------------------------------------------------------------------------------------
constexpr uint32_t ERR_BARRIER = ERR_BARRIER_BIT;
constexpr uint32_t kSpinLimit = 1u << 22;

// Counting barrier: bar[0] is zeroed before each launch and counts arrivals; barrier k of the
// launch is complete when k * nblocks blocks have arrived. Returns true when the launch must end:
// a wait expired (ERR_BARRIER is then set, and every other block leaves at its next wait).
//
// Two forms. Far (blocks on several XCDs, each with its own L2): one release add per block
// (writes this XCD's dirty L2 lines back), relaxed polling, one acquire fence after the wait
// (invalidates stale L2 lines; an acquiring poll would do that on every spin). Near (every block
// on one XCD, so one L2 serves them all): every wave waits until its stores have completed at the
// L2 (wait_stores; __syncthreads alone does not wait for global stores), and every load of data
// written during the launch is a device-scope load (ld_dev) that bypasses the L1, so the add and
// the poll are relaxed and no cache is written back or invalidated. On one XCD this costs ~1.2 us per barrier where the far
// form costs 16-21 us (32 blocks taking turns at one L2's write-back and invalidation).
struct GridBarrier {
    uint32_t* count;
    uint32_t nblocks;
    uint32_t epoch;  // barriers passed (thread 0's copy; every thread's in the flag form)
    uint32_t near;
    uint32_t* flags;  // flag form (near only): block r's epoch at flags[16 r]; null: counting form
    uint32_t rank;
};
// Waits until this wave's global stores have completed (s_waitcnt vmcnt(0); the expcnt and
// lgkmcnt fields of the GFX9 encoding at their maxima wait for nothing). The signal fences keep
// the compiler from moving memory accesses across the wait.
__device__ __forceinline__ void wait_stores() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ __forceinline__ bool barrier_poll(const GridBarrier& g, uint32_t* err) {
    const uint32_t target = g.epoch * g.nblocks;
    uint32_t spins = 0;
    while (__hip_atomic_load(g.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++spins > kSpinLimit) {
            atomicOr(err, ERR_BARRIER);
            return true;
        }
        if ((spins & 255u) == 0u && (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERR_BARRIER))
            return true;
        __builtin_amdgcn_s_sleep(1);
    }
    return false;
}
__device__ __forceinline__ void barrier_add(GridBarrier& g) {
    g.epoch += 1;
    if (g.near) __hip_atomic_fetch_add(g.count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else __hip_atomic_fetch_add(g.count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
// Flag form of the near barrier: each block, once its stores have completed at the L2, stores
// the number of barriers it has reached in its own 64-byte line, and wave 0 of every block reads
// all blocks' lines (one device-scope load per lane) until none is behind. No atomic
// read-modify-write, so the blocks' arrivals are not serialized on one line. The wait is bounded
// as barrier_poll's.
__device__ __forceinline__ void flag_arrive(GridBarrier& g) {
    wait_stores();
    __syncthreads();
    g.epoch += 1;
    if (threadIdx.x == 0) __hip_atomic_store(g.flags + 16u * g.rank, g.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool flag_wait(GridBarrier& g, uint32_t* err) {
    __shared__ uint32_t s_abort_f;
    if (threadIdx.x < 64u) {
        const uint32_t lane = threadIdx.x;
        uint32_t spins = 0, abort = 0;
        while (true) {
            const uint32_t v = lane < g.nblocks
                                   ? __hip_atomic_load(g.flags + 16u * lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                   : g.epoch;
            if (__ballot(v < g.epoch) == 0ull) break;
            if (++spins > kSpinLimit) {
                if (lane == 0u) atomicOr(err, ERR_BARRIER);
                abort = 1u;
                break;
            }
            if ((spins & 255u) == 0u &&
                (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERR_BARRIER)) {
                abort = 1u;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (lane == 0u) s_abort_f = abort;
    }
    __syncthreads();
    return s_abort_f != 0u;
}
__device__ __forceinline__ bool flag_sync(GridBarrier& g, uint32_t* err) {
    flag_arrive(g);
    return flag_wait(g, err);
}
__device__ __forceinline__ bool grid_sync(GridBarrier& g, uint32_t* err) {
    __shared__ uint32_t s_abort;
    if (g.flags) return flag_sync(g, err);
    if (g.near) wait_stores();
    __syncthreads();
    if (threadIdx.x == 0) {
        barrier_add(g);
        const bool abort = barrier_poll(g, err);
        if (!g.near) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_abort = abort ? 1u : 0u;
    }
    __syncthreads();
    return s_abort != 0u;
}
// The same barrier split in two: grid_arrive publishes this block's writes and counts it in;
// grid_wait blocks until every block has arrived. Work between the two (loads of data no other
// block writes in this phase) overlaps the barrier's latency.
__device__ __forceinline__ void grid_arrive(GridBarrier& g) {
    if (g.flags) {
        flag_arrive(g);
        return;
    }
    if (g.near) wait_stores();
    __syncthreads();
    if (threadIdx.x == 0) barrier_add(g);
}
__device__ __forceinline__ bool grid_wait(GridBarrier& g, uint32_t* err) {
    if (g.flags) return flag_wait(g, err);
    __shared__ uint32_t s_abort_w;
    if (threadIdx.x == 0) {
        const bool abort = barrier_poll(g, err);
        if (!g.near) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_abort_w = abort ? 1u : 0u;
    }
    __syncthreads();
    return s_abort_w != 0u;
}
__device__ __forceinline__ GridBarrier grid_barrier_init(uint32_t* bar) {
    GridBarrier g;
    g.count = bar;
    g.nblocks = gridDim.x;
    g.epoch = 0;
    g.near = 0;
    g.flags = nullptr;
    g.rank = 0;
    return g;
}

// The XCD (accelerator complex die) this wave runs on: hardware register XCC_ID, bits 3:0.
__device__ __forceinline__ uint32_t xcc_id() { return __builtin_amdgcn_s_getreg(20 | (3 << 11)) & 15u; }

// Near-barrier set-up of a launch with blocks on every XCD: each block counts itself in
// bar[2 + its XCD] and in bar[1]; once all have, the blocks of the lowest-numbered XCD that has
// any carry on, ranked by their arrival, and the others leave. Returns false for a block that
// leaves (or when the wait expired, with ERR_BARRIER set).
__device__ __forceinline__ bool near_blocks(uint32_t* bar, uint32_t* err, uint32_t& rank, uint32_t& nblocks) {
    __shared__ uint32_t s_go, s_rank, s_n;
    if (threadIdx.x == 0) {
        const uint32_t x = xcc_id();
        const uint32_t r = __hip_atomic_fetch_add(bar + 2 + x, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        GridBarrier reg;
        reg.count = bar + 1;
        reg.nblocks = gridDim.x;
        reg.epoch = 0;
        reg.near = 0;
        barrier_add(reg);
        const bool abort = barrier_poll(reg, err);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        uint32_t chosen = 16u, n = 0u;
        for (uint32_t k = 0; k < 16u && !abort; ++k) {
            const uint32_t c = __hip_atomic_load(bar + 2 + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (c && chosen == 16u) { chosen = k; n = c; }
        }
        s_go = !abort && x == chosen ? 1u : 0u;
        s_rank = r;
        s_n = n;
    }
    __syncthreads();
    rank = s_rank;
    nblocks = s_n;
    return s_go != 0u;
}


struct SolveCfg {
    uint32_t vel_iters, relax_iters, pos_iters;
    uint32_t residual_mode, residual_max;
    float residual_tol;
    uint32_t one_xcd;  // launched on every XCD; the blocks of one carry on with near barriers
    uint32_t ready;    // residual mode: positions split by ready_list (residuals during the tail)
    uint32_t check_every;  // residual mode: the residual after every k-th pass and the last
    uint32_t tiled_lds;    // tiled solve: 1 = groups in LDS where they fit, 0 = always on global memory
};

// A velocity-pass manifold's slot loaded ahead (points in the pass's point order, rows with the
// pass's targets): nothing in it changes during the pass. The two bodies are read after the
// barrier that orders this colour after the last (their inertia loads go out with the velocity
// loads, so holding them across the barrier would only cost registers).
struct PreSlot {
    uint32_t s;  // kNoSlot: none
    uint2 ab;
    float4 nf;
    uint32_t cnt;
    uint32_t idx[4];
    float4 pp[4], row[4], imp[4];
};
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
__device__ __forceinline__ void prefetch_slot(const DevState& d, uint32_t s, uint32_t offset, int use_rest, PreSlot& x) {
    x.s = s;
    x.ab = d.sm_ab[s];
    x.nf = d.sm_nf[s];
    x.cnt = d.sm_cnt[s];
    const uint32_t r = x.cnt ? offset % x.cnt : 0u;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) x.idx[k] = 4u * s + (k < x.cnt ? (k + r) % x.cnt : k);
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        x.pp[k] = x.row[k] = x.imp[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (k < x.cnt) {
            x.pp[k] = d.sp_pos[x.idx[k]];
            x.row[k] = row_for_pass(d.sp_row[x.idx[k]], d.sp_aux[x.idx[k]], use_rest);
            x.imp[k] = ld_dev(d.sp_imp, x.idx[k]);
        }
    }
}
// velocity_manifold for a prefetched slot (same arithmetic, same order)
__device__ __forceinline__ void velocity_prefetched(const DevState& d, PreSlot& x) {
    SBody A = load_sbody(d, x.ab.x), B = load_sbody(d, x.ab.y);
    v3 n = ld3(x.nf);
    v3 t1, t2;
    orthonormal_basis(n, t1, t2);
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k >= x.cnt) break;
        x.imp[k] = velocity_point(A, B, n, t1, t2, x.nf.w, ld3(x.pp[k]), x.row[k], x.imp[k]);
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
        if (k < x.cnt) d.sp_imp[x.idx[k]] = x.imp[k];
    store_sbody(d, x.ab.x, A);
    store_sbody(d, x.ab.y, B);
}

// The trailing colours of at most one wave each (colours [first, C)) are run by one wave, one after
// another: the wave waits for its stores to complete before the next colour's loads, which read
// the L2. That costs each such colour its own latency only, not a grid barrier.
template <class F>
__device__ __forceinline__ void run_tail(const uint32_t* co, uint32_t first, uint32_t C, bool mine, F&& f) {
    if (!mine) return;
    for (uint32_t c = first; c < C; ++c) {
        const uint32_t s = co[c] + (threadIdx.x & 63u);
        if (s < co[c + 1]) f(s);
        wait_stores();
    }
}

// Maximum residual of the solve positions list[j] (j itself without a list) for j in [lo, hi)
// taken by this thread: j = lo + u, lo + u + step, ...; two per iteration, the second index
// clamped rather than branched on, so the loads of both are in flight together (a repeated
// manifold does not change the maximum).
__device__ __forceinline__ float residual_span(const DevState& d, const StepParams& prm, const uint32_t* list, uint32_t lo,
                                               uint32_t hi, uint32_t u, uint32_t step) {
    float r = 0.0f;
    for (uint32_t j = lo + u; j < hi; j += 2u * step) {
        const uint32_t j1 = j + step < hi ? j + step : j;
        const uint32_t s0 = list ? list[j] : j, s1 = list ? list[j1] : j1;
        r = gmax(r, gmax(residual_manifold(d, prm, s0, 0), residual_manifold(d, prm, s1, 0)));
    }
    return r;
}

// Algorithmic bytes of one solve, for the roofline (profiling only; launched after the solve so
// the solve's own time excludes it): per solver manifold with n points, the warm start and each
// velocity or relax pass move 192 + 40 n bytes (both bodies read and written, the points'
// positions, rows and impulses), each position pass 92 + 4 n.
__global__ void k_solve_bytes(DevState d, uint32_t vel_passes, uint32_t relax, uint32_t pos_iters, uint32_t residual_mode) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t passes = residual_mode ? d.cnt->passes : vel_passes;
    unsigned long long bytes = 0;
    if (s < d.cnt->active) {
        const unsigned long long n = d.sm_cnt[s];
        bytes = (192ull + 40ull * n) * (1ull + passes + relax) + pos_iters * (92ull + 4ull * n);
    }
    for (int o = 32; o > 0; o >>= 1) bytes += __shfl_xor(bytes, o, 64);
    if (lane0() && bytes) atomicAdd(&d.cnt->prof_vel_bytes, bytes);
}

// The whole solve of one step (dynamics/solver.h:914-958 after the restitution capture):
// warm start, velocity passes (fixed count or until the residual <= tolerance), relax passes,
// position passes, colour by colour with a grid barrier between colours.
__global__ void __launch_bounds__(kPersistBlock) k_solve_persistent(DevState d, StepParams prm, SolveCfg cfg, uint32_t* bar) {
    GridBarrier g = grid_barrier_init(bar);
    uint32_t* err = &d.cnt->err;
    uint32_t rank = blockIdx.x;
    if (cfg.one_xcd) {
        if (!near_blocks(bar, err, rank, g.nblocks)) return;
        g.near = 1;
        if (g.nblocks <= kBarFlagBlocks) {
            g.flags = bar + kBarFlags;
            g.rank = rank;
        }
    }
    const uint32_t C = d.cnt->colours;
    const uint32_t* co = d.cnt->colour_off;
    // Which thread takes which manifold of a colour does not change the result (the manifolds of
    // one colour write disjoint bodies). Consecutive 64-manifold chunks go to consecutive blocks,
    // so a colour narrower than the grid is spread over every CU instead of filling the first
    // blocks' waves while the others idle.
    const uint32_t tid = ((threadIdx.x >> 6) * g.nblocks + rank) * 64u + (threadIdx.x & 63u);
    const uint32_t nthreads = g.nblocks * blockDim.x;
    const uint32_t T = tail_start(co, C);           // colours [T, C): the tail, one phase
    const uint32_t CP = T + (T < C ? 1u : 0u);      // contact phases per pass
    const bool tail_wave = rank == 0 && threadIdx.x < 64u;
    for (uint32_t c = 0; c < C; ++c) {
        for (uint32_t s = co[c] + tid; s < co[c + 1]; s += nthreads) warm_manifold(d, s);
        if (grid_sync(g, err)) return;
    }
    // joints: warm start, forward layers (solver.h:917)
    const uint32_t L = d.nj_layers;
    const float inv_h = inv_step(prm.h);
    for (uint32_t l = 0; l < L; ++l) {
        uint32_t jb, je;
        const uint32_t* jo;
        joint_layer_range(d, l, false, jb, je, jo);
        for (uint32_t t = jb + tid; t < je; t += nthreads) joint_warm_one(d, jo[t]);
        if (grid_sync(g, err)) return;
    }
    uint32_t passes = 0;
    __shared__ float s_red[kPersistBlock / 64];
    const uint32_t n_active = d.cnt->active;
    if (!cfg.residual_mode) {
        // velocity then relax passes: each pass is the contact colours, then the joint layers
        // (solver.h:939-946, joints alternating direction); the next contact colour's first
        // slot of this thread is loaded between arriving at the barrier and waiting on it
        const uint32_t per = CP + L;
        const uint32_t nph = (cfg.vel_iters + cfg.relax_iters) * per;
        PreSlot x;
        x.s = kNoSlot;
        if (nph && T && co[0] + tid < co[1]) prefetch_slot(d, co[0] + tid, 0u, cfg.vel_iters == 0u ? 1 : 0, x);
        for (uint32_t ph = 0; ph < nph; ++ph) {
            const uint32_t pass = ph / per, k = ph % per;
            const int use_rest = pass >= cfg.vel_iters ? 1 : 0;
            const uint32_t it = use_rest ? pass - cfg.vel_iters : pass;
            if (k < T) {
                const uint32_t c = k;
                if (x.s != kNoSlot) velocity_prefetched(d, x);
                for (uint32_t s = co[c] + tid + nthreads; s < co[c + 1]; s += nthreads) velocity_manifold(d, prm, s, use_rest, it);
            } else if (k < CP) {
                run_tail(co, T, C, tail_wave, [&](uint32_t s) { velocity_manifold(d, prm, s, use_rest, it); });
            } else {
                uint32_t jb, je;
                const uint32_t* jo;
                joint_layer_range(d, k - CP, (it & 1u) != 0u, jb, je, jo);
                for (uint32_t t = jb + tid; t < je; t += nthreads) joint_velocity_one(d, jo[t], inv_h);
            }
            grid_arrive(g);
            x.s = kNoSlot;
            if (ph + 1 < nph) {
                const uint32_t np = (ph + 1) / per, nk = (ph + 1) % per;
                const uint32_t nit = np >= cfg.vel_iters ? np - cfg.vel_iters : np;
                if (nk < T && co[nk] + tid < co[nk + 1]) prefetch_slot(d, co[nk] + tid, nit, np >= cfg.vel_iters ? 1 : 0, x);
            }
            if (grid_wait(g, err)) return;
        }
        if (tid == 0) d.cnt->passes = cfg.vel_iters;
    }
    // residual mode: velocity passes until the residual <= tolerance; as above, each colour's
    // barrier overlaps the load of this thread's first slot of the next colour (of the next pass
    // after the residual's barrier: loads only, harmless when the loop ends there)
    uint32_t max_pass = cfg.residual_mode ? cfg.residual_max : 0u;
    uint32_t checks = 0;  // residual evaluations so far: their parity picks the reduction slot
    PreSlot y;
    y.s = kNoSlot;
    if (max_pass && T && co[0] + tid < co[1]) prefetch_slot(d, co[0] + tid, 0u, 0, y);
    for (uint32_t it = 0; it < max_pass; ++it) {
        for (uint32_t c = 0; c < T; ++c) {
            uint32_t b = co[c], e = co[c + 1];
            if (y.s != kNoSlot) velocity_prefetched(d, y);
            for (uint32_t s = b + tid + nthreads; s < e; s += nthreads) velocity_manifold(d, prm, s, 0, it);
            grid_arrive(g);
            y.s = kNoSlot;
            if (c + 1 < T && co[c + 1] + tid < co[c + 2]) prefetch_slot(d, co[c + 1] + tid, it, 0, y);
            if (grid_wait(g, err)) return;
        }
        // With the split (no joints: joint layers would run after the tail), the blocks the tail
        // leaves idle compute the residual of the manifolds whose bodies no tail colour writes
        // (final before the tail); the rest follow the tail's barrier. The tail's own block stays
        // out of it: residual gathers on its CU would slow the tail's chain of colours.
        float r = 0.0f;
        const bool check = (it + 1u) % cfg.check_every == 0u || it + 1u == max_pass;
        const bool split = check && cfg.ready != 0u && L == 0u && T < C && g.nblocks > 1u;
        if (T < C) {
            if (tail_wave) {
                run_tail(co, T, C, true, [&](uint32_t s) { velocity_manifold(d, prm, s, 0, it); });
            } else if (split && rank != 0u) {
                const uint32_t u = ((threadIdx.x >> 6) * (g.nblocks - 1u) + rank - 1u) * 64u + (threadIdx.x & 63u);
                r = residual_span(d, prm, d.ready_list, 0u, d.ready_cnt[0], u, (g.nblocks - 1u) * blockDim.x);
            }
            if (grid_sync(g, err)) return;
        }
        for (uint32_t l = 0; l < L; ++l) {
            uint32_t jb, je;
            const uint32_t* jo;
            joint_layer_range(d, l, (it & 1u) != 0u, jb, je, jo);
            for (uint32_t t = jb + tid; t < je; t += nthreads) joint_velocity_one(d, jo[t], inv_h);
            if (grid_sync(g, err)) return;
        }
        ++passes;
        if (!check) {
            // no residual this pass: the pass ended at the last colour's, the tail's or the last
            // joint layer's barrier; load the next pass's first slot and go on
            if (T && co[0] + tid < co[1]) prefetch_slot(d, co[0] + tid, it + 1u, 0, y);
            continue;
        }
        r = gmax(r, residual_span(d, prm, split ? d.ready_list : nullptr, split ? d.ready_cnt[0] : 0u, n_active, tid, nthreads));
        for (int o = 32; o > 0; o >>= 1) r = fmaxf(r, __shfl_xor(r, o, 64));
        if (lane0()) s_red[threadIdx.x >> 6] = r;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = 0.0f;
            for (int w = 0; w < kPersistBlock / 64; ++w) m = fmaxf(m, s_red[w]);
            if (m > 0.0f) atomicMax(&d.cnt->res_slot[checks & 1u], __float_as_uint(m));
        }
        grid_arrive(g);
        if (T && co[0] + tid < co[1]) prefetch_slot(d, co[0] + tid, it + 1u, 0, y);
        if (grid_wait(g, err)) return;
        // every block's maximum is in: all threads read the same value. The other slot is
        // cleared for the next evaluation: its last readers read it before this evaluation's
        // barrier, and its next writers add to it only after the next pass's colour barriers.
        float rr = __uint_as_float(__hip_atomic_load(&d.cnt->res_slot[checks & 1u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (tid == 0) {
            d.cnt->residual_bits = __float_as_uint(rr);
            d.cnt->res_slot[(checks + 1u) & 1u] = 0u;
        }
        ++checks;
        if (rr <= cfg.residual_tol) break;
    }
    if (cfg.residual_mode && tid == 0) d.cnt->passes = passes;
    for (uint32_t it = 0; cfg.residual_mode && it < cfg.relax_iters; ++it) {
        for (uint32_t c = 0; c < T; ++c) {
            for (uint32_t s = co[c] + tid; s < co[c + 1]; s += nthreads) velocity_manifold(d, prm, s, 1, it);
            if (grid_sync(g, err)) return;
        }
        if (T < C) {
            run_tail(co, T, C, tail_wave, [&](uint32_t s) { velocity_manifold(d, prm, s, 1, it); });
            if (grid_sync(g, err)) return;
        }
        for (uint32_t l = 0; l < L; ++l) {
            uint32_t jb, je;
            const uint32_t* jo;
            joint_layer_range(d, l, (it & 1u) != 0u, jb, je, jo);
            for (uint32_t t = jb + tid; t < je; t += nthreads) joint_velocity_one(d, jo[t], inv_h);
            if (grid_sync(g, err)) return;
        }
    }
    // solve_position: snapshot, then position passes
    for (uint32_t i = tid; i < d.nb; i += nthreads) {
        float4 p = d.b_pos[i];
        d.sb_x[i] = make_float4(p.x, p.y, p.z, ld_dev(d.sb_v, i).w);
        d.sb_ref[i] = p;
    }
    if (grid_sync(g, err)) return;
    for (uint32_t it = 0; it < cfg.pos_iters; ++it) {
        for (uint32_t c = 0; c < T; ++c) {
            for (uint32_t s = co[c] + tid; s < co[c + 1]; s += nthreads) position_manifold(d, prm, s);
            if (grid_sync(g, err)) return;
        }
        if (T < C) {
            run_tail(co, T, C, tail_wave, [&](uint32_t s) { position_manifold(d, prm, s); });
            if (grid_sync(g, err)) return;
        }
    }
    // one joint position pass after the contacts' (solver.h:954)
    for (uint32_t l = 0; l < L; ++l) {
        uint32_t jb, je;
        const uint32_t* jo;
        joint_layer_range(d, l, false, jb, je, jo);
        for (uint32_t t = jb + tid; t < je; t += nthreads) joint_position_one(d, prm, jo[t]);
        if (grid_sync(g, err)) return;
    }
}

// ---- tiled solve (PXG_ORDER_TILED, set up by pxg_tiles.hip) ----
// One workgroup per tile. A phase's groups share no awake dynamic body, so a workgroup runs its
// group's colours back to back with workgroup barriers, every group of the phase at once; a grid
// barrier separates the three phases. With one group per workgroup and room in LDS, the group's
// bodies (velocities staged in at the phase's start and written out at its end, inertia and
// position once) and the impulses of all the workgroup's solve positions live in LDS; otherwise
// (LDS too small, or fewer workgroups than tiles) the workgroup runs its groups on global memory.
// The arithmetic per manifold is velocity_manifold's / warm_manifold's / residual_manifold's.
constexpr uint32_t kTileLds = 156u * 1024u;  // dynamic LDS per workgroup (one workgroup per CU)
// threads per workgroup: 12 waves (3 per SIMD, within the kernel's register budget), so that more
// colours fit the eight-lane and two-lane plans (431 against 438 ms per residual-mode step on the
// settled 16k pile with 512)
constexpr uint32_t kTileBlock = 768u;

// exclusive scan over the kTileBlock threads of a workgroup
__device__ __forceinline__ uint32_t tile_scan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t wsum[kTileBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= uint32_t(o)) x += y;
    }
    if (lane == 63u) wsum[wid] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < kTileBlock / 64; ++w) {
        if (w < wid) off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

// LDS of one workgroup in the LDS form: per local solve position (the workgroup's colours' ranges
// one after the other) its normal and friction (hn) and (both bodies' indices in the group's table
// as a | b << 16, first point << 3 | point count) (hm); per point (p0) position and the velocity
// passes' normal target, (p1) the three effective masses and the relax passes' normal target, and
// the impulses; per phase the group's table: every body its manifolds touch, with velocity (V: v,
// inverse mass), angular velocity (W: w, 1 if awake dynamic else 0), sb_I's three records (I:
// effective inverse inertia and position) and the body's index (bi). Nothing of a colour's work
// goes to global memory.
struct TLds {
    float4* hn;
    uint2* hm;
    float4 *p0, *p1, *imp;
};
// load_sbody's unpacking of (v, w, sb_I records)
__device__ __forceinline__ SBody tl_body(float4 v, float4 w, float4 a, float4 c, float4 e) {
    SBody x;
    x.v = ld3(v); x.im = v.w;
    x.w = ld3(w); x.act = w.w != 0.0f;
    x.I.c0 = mk(a.x, a.y, a.z); x.I.c1 = mk(a.w, c.x, c.y); x.I.c2 = mk(c.z, c.w, e.x);
    x.x = mk(e.y, e.z, e.w);
    return x;
}
__device__ __forceinline__ SBody tl_entry(const float4* V, const float4* W, const float4* I, uint32_t l) {
    return tl_body(V[l], W[l], I[3u * l], I[3u * l + 1u], I[3u * l + 2u]);
}
// one solve position in LDS: OP 0 = warm_manifold, 1 = velocity_manifold, 2 = velocity_manifold in a
// relax pass (same arithmetic; the targets were selected when the points were staged: the
// velocity passes' is the point row's own, which row_for_pass keeps there)
template <int OP>
__device__ __forceinline__ void tl_run(const TLds& t, float4* V, float4* W, const float4* I, uint32_t L, uint32_t offset) {
    const float4 nf = t.hn[L];
    const uint2 hm = t.hm[L];
    const uint32_t la = hm.x & 0xFFFFu, lb = hm.x >> 16;
    const uint32_t cnt = hm.y & 7u, pt = hm.y >> 3;
    SBody A = tl_entry(V, W, I, la), B = tl_entry(V, W, I, lb);
    v3 n = ld3(nf);
    v3 t1, t2;
    orthonormal_basis(n, t1, t2);
    if (OP == 0) {
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            if (i >= cnt) break;
            const float4 imp = t.imp[pt + i];
            v3 J = n * imp.x + t1 * imp.y + t2 * imp.z;
            if (dot(J, J) > 0.0f) apply_pair(A, B, ld3(t.p0[pt + i]), J);
        }
    } else {
        // point k of the pass is (k + offset % cnt) % cnt, with cnt <= 4 taken without a division
        const uint32_t r = cnt == 3u ? offset - 3u * (__umulhi(offset, 0xAAAAAAABu) >> 1)
                                     : (cnt ? offset & (cnt - 1u) : 0u);
        uint32_t idx[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) idx[k] = k < cnt ? (k + r >= cnt ? k + r - cnt : k + r) : k;
        float4 pp[4], row[4], imp[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            pp[k] = row[k] = imp[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (k < cnt) {
                pp[k] = t.p0[pt + idx[k]];
                const float4 q = t.p1[pt + idx[k]];
                row[k] = make_float4(q.x, q.y, q.z, OP == 2 ? q.w : pp[k].w);
                imp[k] = t.imp[pt + idx[k]];
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            if (k >= cnt) break;
            imp[k] = velocity_point(A, B, n, t1, t2, nf.w, ld3(pp[k]), row[k], imp[k]);
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k)
            if (k < cnt) t.imp[pt + idx[k]] = imp[k];
    }
    if (A.act) { V[la] = st3(A.v, A.im); W[la] = st3(A.w, 1.0f); }
    if (B.act) { V[lb] = st3(B.v, B.im); W[lb] = st3(B.w, 1.0f); }
}
__device__ __forceinline__ void tl_op(int op, const TLds& t, float4* V, float4* W, const float4* I, uint32_t L,
                                      uint32_t offset) {
    if (op == 0) tl_run<0>(t, V, W, I, L, offset);
    else if (op == 1) tl_run<1>(t, V, W, I, L, offset);
    else tl_run<2>(t, V, W, I, L, offset);
}
// ---- two lanes per solve position ----
// A colour lasts as long as its longest manifold's instruction stream on one lane (a wave issues
// one instruction every ~4 cycles whatever the dependences, while further waves on a SIMD cost
// nothing), and half of a contact row is the two bodies' updates, independent of each other. So
// a colour narrow enough takes two adjacent lanes per solve position: lane 0 of the pair owns
// body A, lane 1 body B; both compute the row's shared values from both point velocities (read
// across the pair with a DPP broadcast, bit for bit the same on both lanes), and each applies
// the impulse to its own body with the reference's sign (A: v - J m, B: v + J m, written as
// v + (J m) * -1 / * 1, which IEEE arithmetic makes identical to the subtraction / addition).
__device__ __forceinline__ float pair_lane(float x, int side) {
    const int v = __float_as_int(x);
    // quad_perm [0,0,2,2] or [1,1,3,3]: both lanes of a pair read the pair's lane `side`
    const int r = side ? __builtin_amdgcn_mov_dpp(v, 0xF5, 0xF, 0xF, false) : __builtin_amdgcn_mov_dpp(v, 0xA0, 0xF, 0xF, false);
    return __int_as_float(r);
}
__device__ __forceinline__ v3 pair_lane3(v3 a, int side) {
    return mk(pair_lane(a.x, side), pair_lane(a.y, side), pair_lane(a.z, side));
}
// apply_pair for the lane's own body: sg = -1 for A, +1 for B. x * sg is exact, so the fused
// v + x * sg rounds once, as v - x and v + x do.
__device__ __forceinline__ void apply_own(SBody& s, v3 r, v3 J, float sg) {
    if (!s.act) return;
    const v3 lv = J * s.im, aw = mul(s.I, cross(r, J));
    s.v = mk(__builtin_fmaf(lv.x, sg, s.v.x), __builtin_fmaf(lv.y, sg, s.v.y), __builtin_fmaf(lv.z, sg, s.v.z));
    s.w = mk(__builtin_fmaf(aw.x, sg, s.w.x), __builtin_fmaf(aw.y, sg, s.w.y), __builtin_fmaf(aw.z, sg, s.w.z));
}
// velocity_point with the lane's own body (side 0: A, 1: B)
__device__ __forceinline__ float4 velocity_point2(SBody& S, float sg, v3 n, v3 t1, v3 t2, float mu, v3 pos, float4 row,
                                                  float4 imp) {
    v3 r = pos - S.x;
    {
        const float kn = row.x;
        if (kn > kEps) {
            v3 vme = S.v + cross(S.w, r);
            v3 va = pair_lane3(vme, 0), vb = pair_lane3(vme, 1);
            float vn = dot(vb - va, n);
            float target = row.w;
            float lambda = -(vn - target) / kn;
            float ni = gmax(imp.x + lambda, 0.0f);
            float delta = ni - imp.x;
            imp.x = ni;
            apply_own(S, r, n * delta, sg);
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        v3 dir = t == 0 ? t1 : t2;
        const float kt = t == 0 ? row.y : row.z;
        if (kt <= kEps) continue;
        v3 vme = S.v + cross(S.w, r);
        v3 va = pair_lane3(vme, 0), vb = pair_lane3(vme, 1);
        float vt = dot(vb - va, dir);
        float lt = -vt / kt;
        float maxf = mu * imp.x;
        float old = t == 0 ? imp.y : imp.z;
        float nt = sclamp(old + lt, -maxf, maxf);
        float delta = nt - old;
        if (t == 0) imp.y = nt; else imp.z = nt;
        apply_own(S, r, dir * delta, sg);
    }
    return imp;
}
// tl_run with two lanes per solve position (both lanes of the pair call it with the same L)
template <int OP>
__device__ __forceinline__ void tl_run2(const TLds& t, float4* V, float4* W, const float4* I, uint32_t L, uint32_t offset,
                                        int side) {
    const float4 nf = t.hn[L];
    const uint2 hm = t.hm[L];
    const uint32_t l = side ? hm.x >> 16 : hm.x & 0xFFFFu;
    const uint32_t cnt = hm.y & 7u, pt = hm.y >> 3;
    const float sg = side ? 1.0f : -1.0f;
    SBody S = tl_entry(V, W, I, l);
    v3 n = ld3(nf);
    v3 t1, t2;
    orthonormal_basis(n, t1, t2);
    if (OP == 0) {
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            if (i >= cnt) break;
            const float4 imp = t.imp[pt + i];
            v3 J = n * imp.x + t1 * imp.y + t2 * imp.z;
            if (dot(J, J) > 0.0f) apply_own(S, ld3(t.p0[pt + i]) - S.x, J, sg);
        }
    } else {
        const uint32_t r = cnt == 3u ? offset - 3u * (__umulhi(offset, 0xAAAAAAABu) >> 1)
                                     : (cnt ? offset & (cnt - 1u) : 0u);
        uint32_t idx[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) idx[k] = k < cnt ? (k + r >= cnt ? k + r - cnt : k + r) : k;
        float4 pp[4], row[4], imp[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            pp[k] = row[k] = imp[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (k < cnt) {
                pp[k] = t.p0[pt + idx[k]];
                const float4 q = t.p1[pt + idx[k]];
                row[k] = make_float4(q.x, q.y, q.z, OP == 2 ? q.w : pp[k].w);
                imp[k] = t.imp[pt + idx[k]];
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            if (k >= cnt) break;
            imp[k] = velocity_point2(S, sg, n, t1, t2, nf.w, ld3(pp[k]), row[k], imp[k]);
        }
        if (side == 0) {
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k)
                if (k < cnt) t.imp[pt + idx[k]] = imp[k];
        }
    }
    if (S.act) { V[l] = st3(S.v, S.im); W[l] = st3(S.w, 1.0f); }
}
__device__ __forceinline__ void tl_op2(int op, const TLds& t, float4* V, float4* W, const float4* I, uint32_t L,
                                       uint32_t offset, int side) {
    if (op == 0) tl_run2<0>(t, V, W, I, L, offset, side);
    else if (op == 1) tl_run2<1>(t, V, W, I, L, offset, side);
    else tl_run2<2>(t, V, W, I, L, offset, side);
}
// ---- eight lanes per solve position ----
// A manifold with several points is a long chain of rows; eight lanes cut each row to about half
// the instructions of the two-lane form. Lanes 0-3 of the octet hold body A, lanes 4-7 body B;
// lane c of a quad holds component c of its body's velocities (x, y, z; lane 3 repeats lane 0's
// work and writes nothing). Each lane keeps the other two components of the point's lever arm, the
// row direction and the body's inverse inertia row it needs as constants, so a row's arithmetic is
// per component, with the reference's operations and order:
//   cross(w, r)_c = w_{c+1} r_{c+2} - r_{c+1} w_{c+2}        (the w's from the quad)
//   (vb - va)_c    from the other body's quad                (row_shl / row_shr by 4)
//   dot(d, n)      = (d_x n_x + d_y n_y) + d_z n_z            (d's components from the quad)
//   mul(I, c)_c    = (I_c0 c_x + I_c1 c_y) + I_c2 c_z
// so every lane of the octet computes the same bits as velocity_point. The quad permutes are
// issued where all four lanes of each quad are active (a quad is one body: its lanes share every
// branch), so bound_ctrl (a disabled source reads 0) never applies; it lets the compiler fold
// each permute into the instruction that uses it.
template <int CTRL>
__device__ __forceinline__ float qperm(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, true));
}
constexpr int kQpNext = 1 | (2 << 2) | (0 << 4) | (1 << 6);  // lane c reads lane c + 1 (mod 3)
constexpr int kQpPrev = 2 | (0 << 2) | (1 << 4) | (2 << 6);  // lane c reads lane c + 2 (mod 3)
constexpr int kQpX = 0x00, kQpY = 0x55, kQpZ = 0xAA;
// (vb - va) component c on both quads: body A's lanes take B's value (row_shl 4) minus their own,
// B's lanes their own minus A's (row_shr 4), each instruction writing only its quads (bank mask).
// The leading wait covers the DPP read of u written just before (two wait states after a VALU
// write); the compiler adds the one before the DPP reads of the result.
__device__ __forceinline__ float octet_rel(float u) {
    float d;
    asm("s_nop 1\n\t"
        "v_sub_f32_dpp %0, %1, %1 row_shl:4 row_mask:0xf bank_mask:0x5\n\t"
        "v_subrev_f32_dpp %0, %1, %1 row_shr:4 row_mask:0xf bank_mask:0xa"
        : "=&v"(d)
        : "v"(u));
    return d;
}
__device__ __forceinline__ float rot3(v3 a, uint32_t c) { return c == 0u ? a.x : (c == 1u ? a.y : a.z); }
// a row on the octet takes the row direction rotated to the lane (D0 = dir_c, D1 = dir_{c+1},
// D2 = dir_{c+2}), the lever arm's other two components (R1 = r_{c+1}, R2 = r_{c+2}) and the
// body's inverse inertia row c (M0, M1, M2); OctRow is the lane's velocity components
struct OctRow {
    float v, w;
};
// the relative velocity along dir (full vector, the same on every lane): dot(vb - va, dir) with
// the three components of vb - va read from the quad
__device__ __forceinline__ float octet_vrel(const OctRow& s, v3 dir, float R1, float R2) {
    const float u = s.v + (qperm<kQpNext>(s.w) * R2 - R1 * qperm<kQpPrev>(s.w));
    const float d = octet_rel(u);
    return (qperm<kQpX>(d) * dir.x + qperm<kQpY>(d) * dir.y) + qperm<kQpZ>(d) * dir.z;
}
// (computed on every lane, kept where `on`: no branch in the row's chain)
__device__ __forceinline__ void octet_apply(OctRow& s, bool on, float im, float sg, float D0, float D1, float D2,
                                            float R1, float R2, float M0, float M1, float M2, float delta) {
    const float J0 = D0 * delta, J1 = D1 * delta, J2 = D2 * delta;
    const float cr = R1 * J2 - J1 * R2;
    const float aw = (M0 * qperm<kQpX>(cr) + M1 * qperm<kQpY>(cr)) + M2 * qperm<kQpZ>(cr);
    const float nv = __builtin_fmaf(J0 * im, sg, s.v), nw = __builtin_fmaf(aw, sg, s.w);
    s.v = on ? nv : s.v;
    s.w = on ? nw : s.w;
}
// velocity_manifold (a velocity pass, or a relax pass when relax) of local solve position L on
// the octet; lane = the lane's index in the octet
__device__ __forceinline__ void tl_run8(const TLds& t, float4* V, float4* W, const float4* I, uint32_t L,
                                        uint32_t offset, uint32_t lane, bool relax) {
    const float4 nf = t.hn[L];
    const uint2 hm = t.hm[L];
    const uint32_t side = lane >> 2, c = lane & 3u, cc = c == 3u ? 0u : c;
    const uint32_t c1 = cc == 2u ? 0u : cc + 1u, c2 = cc == 0u ? 2u : cc - 1u;
    const uint32_t l = side ? hm.x >> 16 : hm.x & 0xFFFFu;
    const uint32_t cnt = hm.y & 7u, pt = hm.y >> 3;
    const float sg = side ? 1.0f : -1.0f;
    const float* vb = reinterpret_cast<const float*>(V + l);
    const float* wb = reinterpret_cast<const float*>(W + l);
    const float* ib = reinterpret_cast<const float*>(I + 3u * l);
    OctRow s;
    s.v = vb[cc];
    s.w = wb[cc];
    const float im = vb[3];
    const bool act = wb[3] != 0.0f;
    const float M0 = ib[cc], M1 = ib[3u + cc], M2 = ib[6u + cc];
    const float X1 = ib[9u + c1], X2 = ib[9u + c2];
    const v3 n = ld3(nf);
    v3 t1, t2;
    orthonormal_basis(n, t1, t2);
    const float N0 = rot3(n, cc), N1 = rot3(n, c1), N2 = rot3(n, c2);
    const float A0 = rot3(t1, cc), A1 = rot3(t1, c1), A2 = rot3(t1, c2);
    const float B0 = rot3(t2, cc), B1 = rot3(t2, c1), B2 = rot3(t2, c2);
    const uint32_t r = cnt == 3u ? offset - 3u * (__umulhi(offset, 0xAAAAAAABu) >> 1)
                                 : (cnt ? offset & (cnt - 1u) : 0u);
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k >= cnt) break;
        const uint32_t q = pt + (k + r >= cnt ? k + r - cnt : k + r);
        const float* pp = reinterpret_cast<const float*>(t.p0 + q);
        const float4 row = t.p1[q];
        float4 imp = t.imp[q];
        const float R1 = pp[c1] - X1, R2 = pp[c2] - X2;
        // the three rows without branches: a row whose effective mass is not above kEps (the
        // reference skips it) is computed and its results dropped
        {
            const bool on = row.x > kEps;
            const float target = relax ? row.w : pp[3];
            const float vn = octet_vrel(s, n, R1, R2);
            const float lambda = -(vn - target) / row.x;
            const float ni = gmax(imp.x + lambda, 0.0f);
            const float delta = ni - imp.x;
            imp.x = on ? ni : imp.x;
            octet_apply(s, act && on, im, sg, N0, N1, N2, R1, R2, M0, M1, M2, delta);
        }
        {
            const bool on = !(row.y <= kEps);
            const float vt = octet_vrel(s, t1, R1, R2);
            const float lt = -vt / row.y;
            const float maxf = nf.w * imp.x;
            const float nt = sclamp(imp.y + lt, -maxf, maxf);
            const float delta = nt - imp.y;
            imp.y = on ? nt : imp.y;
            octet_apply(s, act && on, im, sg, A0, A1, A2, R1, R2, M0, M1, M2, delta);
        }
        {
            const bool on = !(row.z <= kEps);
            const float vt = octet_vrel(s, t2, R1, R2);
            const float lt = -vt / row.z;
            const float maxf = nf.w * imp.x;
            const float nt = sclamp(imp.z + lt, -maxf, maxf);
            const float delta = nt - imp.z;
            imp.z = on ? nt : imp.z;
            octet_apply(s, act && on, im, sg, B0, B1, B2, R1, R2, M0, M1, M2, delta);
        }
        if (lane == 0u) t.imp[q] = imp;
    }
    if (act && c < 3u) {
        reinterpret_cast<float*>(V + l)[c] = s.v;
        reinterpret_cast<float*>(W + l)[c] = s.w;
    }
}
// residual_manifold (velocity passes) of a local solve position, all from LDS (the group's
// velocities freshly staged after the pass). The normal target -sep / h of a separated point is
// the velocity passes' target staged in p0.w.
__device__ __forceinline__ float tl_residual(const TLds& t, const float4* V, const float4* W, const float4* I, uint32_t L) {
    float r = 0.0f;
    const float4 nf = t.hn[L];
    const uint2 hm = t.hm[L];
    SBody A = tl_entry(V, W, I, hm.x & 0xFFFFu), B = tl_entry(V, W, I, hm.x >> 16);
    v3 n = ld3(nf);
    v3 t1, t2;
    orthonormal_basis(n, t1, t2);
    const uint32_t cnt = hm.y & 7u, pt = hm.y >> 3;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        if (i >= cnt) break;
        const float4 pp = t.p0[pt + i], imp = t.imp[pt + i];
        v3 pos = ld3(pp);
        v3 ra = pos - A.x, rb = pos - B.x;
        v3 dv = (B.v + cross(B.w, rb)) - (A.v + cross(A.w, ra));
        float vn = dot(dv, n);
        float target = pp.w;
        float rn = imp.x > 0.0f ? fabsf(vn - target) : gmax(0.0f, target - vn);
        r = gmax(r, rn);
        float maxf = nf.w * imp.x;
        for (int q = 0; q < 2; ++q) {
            float vt = dot(dv, q == 0 ? t1 : t2);
            float lt = q == 0 ? imp.y : imp.z;
            float rt = 0.0f;
            if (maxf > 0.0f) rt = fabsf(lt) < maxf ? fabsf(vt) : gmax(0.0f, lt < 0.0f ? -vt : vt);
            r = gmax(r, rt);
        }
    }
    return r;
}

// per-workgroup state of the tiled solve (shared)
struct TileBlock {
    uint32_t cols[3], base[3];       // colours per phase, first global colour of each
    uint32_t lds;                    // 1: the LDS form
    uint32_t nslots, npts;           // local solve positions and their points
    uint32_t lo[3 * kTileColours], hi[3 * kTileColours], lb[3 * kTileColours];  // per (phase, colour)
    uint32_t nbody[3], body0[3];     // the group's body table
    uint32_t p_off, v_off[3];        // LDS float4 offsets: points, each phase's table
    uint16_t plan[3 * kTileColours]; // per (phase, colour): octets and the lanes of the rest
};
// How a colour's velocity and relax passes spread over the workgroup's lanes (its solve positions
// are ordered by point count, most first): the first H on eight lanes each, the rest on two lanes
// or one each, all in one round; or, when that does not fit, every position on two lanes (at most
// half as many positions as lanes) or on one lane in rounds. H is every position when eight lanes
// each fit, else the positions with two or more points rounded up to a multiple of 8 (whole
// waves). plan = H | mode << 8.
constexpr uint32_t kPlanOct = 0u, kPlanMix2 = 1u, kPlanMix1 = 2u, kPlanOld = 3u;
__device__ __forceinline__ uint32_t tile_plan(uint32_t n, uint32_t heavy, uint32_t nth) {
    if (n * 8u <= nth) return n | (kPlanOct << 8);
    const uint32_t H = (heavy + 7u) & ~7u;
    if (H * 8u <= nth && H <= n) {
        const uint32_t lanes = nth - H * 8u, rest = n - H;
        if (rest * 2u <= lanes) return H | (kPlanMix2 << 8);
        if (rest <= lanes) return H | (kPlanMix1 << 8);
    }
    return kPlanOld << 8;
}
// the (phase, colour) entry holding local solve position L: the last with lb <= L
__device__ __forceinline__ uint32_t tl_find(const TileBlock& tb, uint32_t L) {
    uint32_t lo = 0, hi = 3u * kTileColours;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tb.lb[mid] <= L) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kTileBlock) k_solve_tiled(DevState d, StepParams prm, SolveCfg cfg, uint32_t* bar) {
    extern __shared__ float4 t_lds[];
    __shared__ TileBlock tb;
    __shared__ float s_red[kTileBlock / 64];
    GridBarrier g = grid_barrier_init(bar);
    uint32_t* err = &d.cnt->err;
    uint32_t rank = blockIdx.x;
    if (cfg.one_xcd) {
        if (!near_blocks(bar, err, rank, g.nblocks)) return;
        g.near = 1;
        if (g.nblocks <= kBarFlagBlocks) {
            g.flags = bar + kBarFlags;
            g.rank = rank;
        }
    }
    const uint32_t nblk = g.nblocks, tid = threadIdx.x, nth = kTileBlock;
    const uint32_t gtid = rank * nth + tid, gnth = nblk * nth;
    // ---- set-up: this workgroup's groups, colour ranges, and whether they fit in LDS ----
    const bool one_group = nblk == kTiles;
    if (tid == 0) {
        // at most kTileColours in all (more is a colouring error, already reported: the excess
        // colours are left out rather than read past t_off)
        const uint32_t* c = d.cnt->tile_cols;
        const uint32_t c0 = c[0] < kTileColours ? c[0] : kTileColours;
        const uint32_t c1 = c[1] < kTileColours - c0 ? c[1] : kTileColours - c0;
        const uint32_t c2 = c[2] < kTileColours - c0 - c1 ? c[2] : kTileColours - c0 - c1;
        tb.cols[0] = c0; tb.cols[1] = c1; tb.cols[2] = c2;
        tb.base[0] = 0u; tb.base[1] = c0; tb.base[2] = c0 + c1;
        tb.nslots = 0u;
    }
    __syncthreads();
    // ranges of this workgroup's colours (LDS form: group = rank in phases 0 and 1, group 0 of
    // phase 2 on workgroup 0)
    if (one_group) {
        for (uint32_t k = tid; k < 3u * kTileColours; k += nth) {
            const uint32_t ph = k / kTileColours, c = k - ph * kTileColours;
            const bool own = ph < 2u || rank == 0u;
            uint32_t lo = 0u, hi = 0u;
            if (own && c < tb.cols[ph]) {
                const uint32_t b = (tb.base[ph] + c) * kTiles + (ph < 2u ? rank : 0u);
                lo = d.t_off[b];
                hi = d.t_off[b + 1u];
            }
            tb.lo[k] = lo;
            tb.hi[k] = hi;
        }
        if (tid < 3u) {
            const uint32_t ph = tid;
            const bool own = ph < 2u || rank == 0u;
            const uint32_t key = ph * kTiles + (ph < 2u ? rank : 0u);
            tb.body0[ph] = own ? d.tb_off[key] : 0u;
            tb.nbody[ph] = own ? d.tb_off[key + 1u] - d.tb_off[key] : 0u;
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t n = 0;
            for (uint32_t k = 0; k < 3u * kTileColours; ++k) {
                tb.lb[k] = n;
                n += tb.hi[k] - tb.lo[k];
            }
            tb.nslots = n;
        }
        __syncthreads();
    }
    TLds t;
    const uint32_t hm_f4 = (tb.nslots + 1u) / 2u;  // float4s of the hm array
    t.hn = t_lds;
    t.hm = reinterpret_cast<uint2*>(t_lds + tb.nslots);
    bool lds = one_group && cfg.tiled_lds && (tb.nslots + hm_f4) * 16u <= kTileLds;
    if (lds) {
        // headers, with each position's first point from a scan of the point counts
        uint32_t carry = 0;
        for (uint32_t base = 0; base < tb.nslots; base += nth) {
            const uint32_t L = base + tid;
            uint32_t cnt = 0, s = 0;
            if (L < tb.nslots) {
                const uint32_t k = tl_find(tb, L);
                s = tb.lo[k] + (L - tb.lb[k]);
                cnt = d.sm_cnt[s];
            }
            uint32_t tot;
            const uint32_t ex = tile_scan(cnt, tot);
            if (L < tb.nslots) {
                t.hn[L] = d.sm_nf[s];
                t.hm[L] = make_uint2(d.sm_loc[s], ((carry + ex) << 3) | cnt);
            }
            carry += tot;
        }
        if (tid == 0) {
            tb.npts = carry;
            uint32_t off = tb.nslots + hm_f4;
            tb.p_off = off;
            off += 3u * carry;
            for (uint32_t ph = 0; ph < 3u; ++ph) {
                tb.v_off[ph] = off;
                off += 5u * tb.nbody[ph] + (tb.nbody[ph] + 3u) / 4u;  // V, W, 3 I, then the indices
            }
            tb.lds = off * 16u <= kTileLds ? 1u : 0u;
            atomicMax(&d.cnt->tile_need, off * 16u);
        }
        __syncthreads();
        lds = tb.lds != 0u;
    }
    if (tid == 0) atomicAdd(&d.cnt->tile_wg, 1u | (lds ? 0x100u : 0u));
    if (lds) {
        // each colour's lane plan from its positions' point counts (more points first)
        for (uint32_t k = tid; k < 3u * kTileColours; k += nth) {
            const uint32_t n = tb.hi[k] - tb.lo[k], L0 = tb.lb[k];
            uint32_t a = 0, b = n;
            while (a < b) {
                const uint32_t mid = (a + b) >> 1;
                if ((t.hm[L0 + mid].y & 7u) >= 2u) a = mid + 1u;
                else b = mid;
            }
            tb.plan[k] = uint16_t(tile_plan(n, a, nth));
        }
        t.p0 = t_lds + tb.p_off;
        t.p1 = t.p0 + tb.npts;
        t.imp = t.p1 + tb.npts;
        // the points: positions, targets, effective masses, impulses
        for (uint32_t L = tid; L < tb.nslots; L += nth) {
            const uint32_t k = tl_find(tb, L), s = tb.lo[k] + (L - tb.lb[k]);
            const uint2 hm = t.hm[L];
            const uint32_t cnt = hm.y & 7u, pt = hm.y >> 3;
            for (uint32_t q = 0; q < cnt; ++q) {
                const float4 pp = d.sp_pos[4u * s + q], row = d.sp_row[4u * s + q];
                const float4 rr = row_for_pass(row, d.sp_aux[4u * s + q], 1);
                t.p0[pt + q] = make_float4(pp.x, pp.y, pp.z, row.w);
                t.p1[pt + q] = make_float4(row.x, row.y, row.z, rr.w);
                t.imp[pt + q] = d.sp_imp[4u * s + q];
            }
        }
        // the tables' constants: inertia, position, body index
        for (uint32_t ph = 0; ph < 3u; ++ph) {
            const uint32_t n = tb.nbody[ph];
            float4* I = t_lds + tb.v_off[ph] + 2u * n;
            uint32_t* bi = reinterpret_cast<uint32_t*>(I + 3u * n);
            for (uint32_t i = tid; i < n; i += nth) {
                const uint32_t b = d.tb_list[tb.body0[ph] + i];
                bi[i] = b;
                I[3u * i] = d.sb_I[3u * b];
                I[3u * i + 1u] = d.sb_I[3u * b + 1u];
                I[3u * i + 2u] = d.sb_I[3u * b + 2u];
            }
        }
        __syncthreads();
    }
    // stages a phase's table velocities in from global memory (every entry: those of bodies that
    // are not awake dynamic never change, but a table is small)
    auto stage_in = [&](uint32_t ph) __attribute__((always_inline)) {
        const uint32_t n = tb.nbody[ph];
        float4* V = t_lds + tb.v_off[ph];
        const uint32_t* bi = reinterpret_cast<const uint32_t*>(V + 5u * n);
        for (uint32_t i = tid; i < n; i += nth) {
            const uint32_t b = bi[i];
            V[i] = ld_dev(d.sb_v, b);
            V[n + i] = ld_dev(d.sb_w, b);
        }
    };
    // profiling: workgroup 0's time in its colours, staging, grid barriers and residuals (10 ns
    // ticks of the real-time counter), added to Counters::prof_stage at the end
    const bool prof = prm.profile != 0u && gtid == 0u;
    unsigned long long tk_col = 0, tk_stage = 0, tk_bar = 0, tk_res = 0;
    auto tick = [&]() __attribute__((always_inline)) -> unsigned long long {
        return prof ? static_cast<unsigned long long>(wall_clock64()) : 0ull;
    };
    // ---- one phase of a pass: op 0 warm start, 1 velocity, 2 relax ----
    auto phase = [&](uint32_t ph, int op, uint32_t offset) __attribute__((always_inline)) {
        if (lds) {
            if (ph == 2u && rank != 0u) return;
            const unsigned long long t0 = tick();
            const uint32_t n = tb.nbody[ph], C = tb.cols[ph], k0 = ph * kTileColours;
            float4* V = t_lds + tb.v_off[ph];
            float4* W = V + n;
            const float4* I = W + n;
            const uint32_t* bi = reinterpret_cast<const uint32_t*>(I + 3u * n);
            stage_in(ph);
            __syncthreads();
            const unsigned long long t1 = tick();
            tk_stage += t1 - t0;
            for (uint32_t c = 0; c < C; ++c) {
                const uint32_t k = k0 + c, lo = tb.lo[k], hi = tb.hi[k], lb = tb.lb[k], n = hi - lo;
                const uint32_t plan = tb.plan[k];
                uint32_t H = plan & 0xFFu, mode = plan >> 8;
                // the warm start and colours without a plan: two lanes per solve position when the
                // whole colour fits in one round, else one lane in rounds
                if (op == 0 || mode == kPlanOld) {
                    H = 0u;
                    mode = n <= nth / 2u ? kPlanMix2 : kPlanMix1;
                }
                // eight lanes for the first H positions, two or one for the rest (whole waves
                // each: H is all positions or a multiple of 8)
                if (tid < 8u * H) {
                    tl_run8(t, V, W, I, lb + (tid >> 3), offset, tid & 7u, op == 2);
                } else if (mode == kPlanMix2) {
                    const uint32_t j = H + ((tid - 8u * H) >> 1);
                    if (j < n) tl_op2(op, t, V, W, I, lb + j, offset, int(tid & 1u));
                } else if (mode == kPlanMix1) {
                    for (uint32_t j = H + (tid - 8u * H); j < n; j += nth) tl_op(op, t, V, W, I, lb + j, offset);
                }
                __syncthreads();
            }
            const unsigned long long t2 = tick();
            tk_col += t2 - t1;
            for (uint32_t i = tid; i < n; i += nth) {
                const float4 w = W[i];
                if (w.w != 0.0f) {
                    const uint32_t b = bi[i];
                    d.sb_v[b] = V[i];
                    d.sb_w[b] = w;
                }
            }
            tk_stage += tick() - t2;
            return;
        }
        const uint32_t g0 = ph < 2u ? rank : (rank == 0u ? 0u : 1u), gend = ph < 2u ? kTiles : 1u;
        for (uint32_t gg = g0; gg < gend; gg += nblk) {
            for (uint32_t c = 0; c < tb.cols[ph]; ++c) {
                const uint32_t b = (tb.base[ph] + c) * kTiles + gg;
                const uint32_t lo = d.t_off[b], hi = d.t_off[b + 1u];
                for (uint32_t s = lo + tid; s < hi; s += nth) {
                    if (op == 0) warm_manifold(d, s);
                    else velocity_manifold(d, prm, s, op == 2 ? 1 : 0, offset);
                }
                wait_stores();
                __syncthreads();
            }
        }
    };
    // a whole pass: the three phases, a grid barrier after each that has colours
    auto pass = [&](int op, uint32_t offset) __attribute__((always_inline)) -> bool {
        for (uint32_t ph = 0; ph < 3u; ++ph) {
            if (!tb.cols[ph]) continue;
            phase(ph, op, offset);
            const unsigned long long t0 = tick();
            if (grid_sync(g, err)) return true;
            tk_bar += tick() - t0;
        }
        return false;
    };
    // maximum residual over this workgroup's solve positions (after the pass's last grid barrier)
    auto block_residual = [&]() __attribute__((always_inline)) -> float {
        float r = 0.0f;
        if (lds) {
            for (uint32_t ph = 0; ph < 3u; ++ph) stage_in(ph);
            __syncthreads();
            for (uint32_t ph = 0; ph < 3u; ++ph) {
                const uint32_t n = tb.nbody[ph];
                const float4* V = t_lds + tb.v_off[ph];
                const float4* W = V + n;
                const float4* I = W + n;
                const uint32_t k0 = ph * kTileColours;
                const uint32_t L0 = tb.lb[k0], L1 = ph < 2u ? tb.lb[k0 + kTileColours] : tb.nslots;
                for (uint32_t L = L0 + tid; L < L1; L += nth) r = gmax(r, tl_residual(t, V, W, I, L));
            }
        } else {
            for (uint32_t ph = 0; ph < 3u; ++ph) {
                const uint32_t g0 = ph < 2u ? rank : (rank == 0u ? 0u : 1u), gend = ph < 2u ? kTiles : 1u;
                for (uint32_t gg = g0; gg < gend; gg += nblk)
                    for (uint32_t c = 0; c < tb.cols[ph]; ++c) {
                        const uint32_t b = (tb.base[ph] + c) * kTiles + gg;
                        const uint32_t lo = d.t_off[b], hi = d.t_off[b + 1u];
                        for (uint32_t s = lo + tid; s < hi; s += nth) r = gmax(r, residual_manifold(d, prm, s, 0));
                    }
            }
        }
        return r;
    };
    // ---- the solve (dynamics/solver.h:914-958 after the restitution capture) ----
    if (pass(0, 0u)) return;
    uint32_t passes = 0;
    if (!cfg.residual_mode) {
        for (uint32_t it = 0; it < cfg.vel_iters; ++it)
            if (pass(1, it)) return;
        passes = cfg.vel_iters;
    }
    const uint32_t max_pass = cfg.residual_mode ? cfg.residual_max : 0u;
    uint32_t checks = 0;
    for (uint32_t it = 0; it < max_pass; ++it) {
        if (pass(1, it)) return;
        ++passes;
        const bool check = (it + 1u) % cfg.check_every == 0u || it + 1u == max_pass;
        if (!check) continue;
        const unsigned long long tr = tick();
        float r = block_residual();
        for (int o = 32; o > 0; o >>= 1) r = fmaxf(r, __shfl_xor(r, o, 64));
        if (lane0()) s_red[tid >> 6] = r;
        __syncthreads();
        if (tid == 0) {
            float m = 0.0f;
            for (uint32_t w = 0; w < kTileBlock / 64; ++w) m = fmaxf(m, s_red[w]);
            if (m > 0.0f) atomicMax(&d.cnt->res_slot[checks & 1u], __float_as_uint(m));
        }
        if (grid_sync(g, err)) return;
        // every workgroup's maximum is in; the other slot is cleared for the next evaluation (its
        // last readers read it before this evaluation's barrier, its next writers add to it after
        // the next pass's barriers)
        const float rr = __uint_as_float(__hip_atomic_load(&d.cnt->res_slot[checks & 1u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (gtid == 0) {
            d.cnt->residual_bits = __float_as_uint(rr);
            d.cnt->res_slot[(checks + 1u) & 1u] = 0u;
        }
        ++checks;
        tk_res += tick() - tr;
        if (rr <= cfg.residual_tol) break;
    }
    if (gtid == 0) d.cnt->passes = passes;
    for (uint32_t it = 0; it < cfg.relax_iters; ++it)
        if (pass(2, it)) return;
    if (prof) {
        atomicAdd(&d.cnt->prof_stage[0], tk_col);
        atomicAdd(&d.cnt->prof_stage[1], tk_stage);
        atomicAdd(&d.cnt->prof_stage[2], tk_bar);
        atomicAdd(&d.cnt->prof_stage[3], tk_res);
    }
    if (lds) {
        // the impulses back for the warm-start cache
        for (uint32_t L = tid; L < tb.nslots; L += nth) {
            const uint32_t k = tl_find(tb, L), s = tb.lo[k] + (L - tb.lb[k]);
            const uint2 hm = t.hm[L];
            const uint32_t cnt = hm.y & 7u, pt = hm.y >> 3;
            for (uint32_t q = 0; q < cnt; ++q) d.sp_imp[4u * s + q] = t.imp[pt + q];
        }
    }
    // solve_position: snapshot, then position passes, on global memory
    for (uint32_t i = gtid; i < d.nb; i += gnth) {
        const float4 p = d.b_pos[i];
        d.sb_x[i] = make_float4(p.x, p.y, p.z, ld_dev(d.sb_v, i).w);
        d.sb_ref[i] = p;
    }
    if (grid_sync(g, err)) return;
    for (uint32_t it = 0; it < cfg.pos_iters; ++it) {
        for (uint32_t ph = 0; ph < 3u; ++ph) {
            if (!tb.cols[ph]) continue;
            const uint32_t g0 = ph < 2u ? rank : (rank == 0u ? 0u : 1u), gend = ph < 2u ? kTiles : 1u;
            for (uint32_t gg = g0; gg < gend; gg += nblk)
                for (uint32_t c = 0; c < tb.cols[ph]; ++c) {
                    const uint32_t b = (tb.base[ph] + c) * kTiles + gg;
                    const uint32_t lo = d.t_off[b], hi = d.t_off[b + 1u];
                    for (uint32_t s = lo + tid; s < hi; s += nth) position_manifold(d, prm, s);
                    wait_stores();
                    __syncthreads();
                }
            if (grid_sync(g, err)) return;
        }
    }
}

// ---- reference solve order ----
// The reference solves manifolds one after the other in pair order (broadphase/pair_cache.h:23-52,
// dynamics/solver.h:212-281). Two manifolds that write no common body commute, so any order that
// keeps every awake dynamic body's manifolds in pair order gives the same result. Each such body's
// manifolds form a chain in pair order; a manifold's colour is one more than the larger colour
// of its two predecessors (its chain layer), so colour-major order keeps every chain in order.
__device__ __forceinline__ unsigned long long pair_key(const DevState& d, uint32_t m) {
    uint2 sh = d.m_shapes[m];
    return (static_cast<unsigned long long>(sh.x) << 32) | sh.y;
}
__device__ __forceinline__ void ref_count(const DevState& d, uint32_t m) {
    d.m_pred[2 * m] = kNoPred;
    d.m_pred[2 * m + 1] = kNoPred;
    if (d.m_colour[m] != kUncoloured) return;
    uint2 ab = d.m_ab[m];
    if (body_active(d.b_flags[ab.x])) atomicAdd(&d.adj_cnt[ab.x], 1u);
    if (body_active(d.b_flags[ab.y])) atomicAdd(&d.adj_cnt[ab.y], 1u);
}
__device__ __forceinline__ void ref_scatter(const DevState& d, uint32_t m) {
    if (d.m_colour[m] != kUncoloured) return;
    uint2 ab = d.m_ab[m];
    if (body_active(d.b_flags[ab.x])) d.adj[d.adj_start[ab.x] + atomicAdd(&d.adj_cur[ab.x], 1u)] = m;
    if (body_active(d.b_flags[ab.y])) d.adj[d.adj_start[ab.y] + atomicAdd(&d.adj_cur[ab.y], 1u)] = m;
}
// sorts body x's chain into pair order and links each manifold to its predecessor on x
__device__ __forceinline__ void ref_chain(const DevState& d, uint32_t x) {
    const uint32_t n = d.adj_cnt[x];
    if (n == 0u) return;
    uint32_t* seg = d.adj + d.adj_start[x];
    for (uint32_t a = 1; a < n; ++a) {
        uint32_t v = seg[a];
        unsigned long long kv = pair_key(d, v);
        uint32_t b = a;
        while (b > 0 && pair_key(d, seg[b - 1]) > kv) {
            seg[b] = seg[b - 1];
            --b;
        }
        seg[b] = v;
    }
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t m = seg[k];
        uint32_t side = d.m_ab[m].x == x ? 0u : 1u;
        d.m_pred[2 * m + side] = k ? seg[k - 1] : kNoPred;
    }
}
// one layering round for manifold m: coloured once both predecessors are; returns the colour + 1
// (0: not coloured this round)
__device__ __forceinline__ uint32_t ref_layer(const DevState& d, uint32_t m, uint32_t* err) {
    if (d.m_colour[m] != kUncoloured) return 0u;
    const uint32_t pa = d.m_pred[2 * m], pb = d.m_pred[2 * m + 1];
    uint32_t c = 0u;
    if (pa != kNoPred) {
        uint32_t ca = __hip_atomic_load(&d.m_colour[pa], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (ca >= uint32_t(kMaxColours)) return 0u;
        c = ca + 1u;
    }
    if (pb != kNoPred) {
        uint32_t cb = __hip_atomic_load(&d.m_colour[pb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cb >= uint32_t(kMaxColours)) return 0u;
        c = c > cb + 1u ? c : cb + 1u;
    }
    if (c >= uint32_t(kMaxColours)) {
        atomicOr(err, ERR_COLOURS);
        return 0u;
    }
    __hip_atomic_store(&d.m_colour[m], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return c + 1u;
}
// wave-aggregated bookkeeping of a colouring step: `top` = colour + 1 of a manifold coloured by
// this lane (0: none)
__device__ __forceinline__ void colour_taken(const DevState& d, uint32_t top) {
    uint32_t took = wave_sum(top ? 1u : 0u);
    top = wave_max(top);
    if (lane0() && took) {
        atomicSub(&d.cnt->remaining, took);
        atomicMax(&d.cnt->colours, top);
    }
}
__global__ void k_ref_count(DevState d, uint32_t bound) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < bound && m < d.cnt->manifolds) ref_count(d, m);
}
__global__ void k_ref_scatter(DevState d, uint32_t bound) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < bound && m < d.cnt->manifolds) ref_scatter(d, m);
}
__global__ void k_ref_chain(DevState d) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < d.nb) ref_chain(d, x);
}
__global__ void k_ref_layer(DevState d, uint32_t bound) {
    if (d.cnt->remaining == 0u) return;
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t top = (m < bound && m < d.cnt->manifolds) ? ref_layer(d, m, &d.cnt->err) : 0u;
    colour_taken(d, top);
}
// exclusive scan of n counts by one block (any block size up to 1024)
__device__ void block_scan_range(const uint32_t* in, uint32_t* out, uint32_t n) {
    __shared__ uint32_t wtot[16];
    __shared__ uint32_t carry;
    const uint32_t t = threadIdx.x, nt = blockDim.x, lane = t & 63u, wid = t >> 6, nw = (nt + 63u) >> 6;
    if (t == 0) carry = 0u;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += nt) {
        uint32_t i = base + t;
        uint32_t v = i < n ? in[i] : 0u, x = v;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if (lane >= uint32_t(o)) x += y;
        }
        if (lane == 63u) wtot[wid] = x;
        __syncthreads();
        uint32_t off = carry, tot = 0u;
        for (uint32_t w = 0; w < nw; ++w) {
            if (w < wid) off += wtot[w];
            tot += wtot[w];
        }
        if (i < n) out[i] = off + x - v;
        __syncthreads();
        if (t == 0) carry += tot;
        __syncthreads();
    }
}

// The whole colouring in one cooperative launch. Contacts persist from step to step, so each
// active manifold first tries the colour its first contact point had last step (kept in the
// warm-start entry). A hinted colour is kept unless another hinted manifold claims the same colour
// at the same awake dynamic body, in which case every claimant drops it; the outcome does not
// depend on timing. Jones-Plassmann rounds (first-fit over the per-body masks) colour the rest.
// reference: 0 = coloured order starting from last step's colours, 1 = the reference order's
// chain layers, 2 = coloured order from scratch (Jones-Plassmann rounds only), 3 = the tiled
// order's colours: Jones-Plassmann rounds with one mask word per (phase, body) (col_best and
// col_mask then hold 3 * nb words, phase-major), each phase's colours counted in tile_cols.
__global__ void __launch_bounds__(kPersistBlock) k_colour_persistent(DevState d, uint32_t max_rounds, uint32_t reference,
                                                                   uint32_t* bar) {
    GridBarrier g = grid_barrier_init(bar);
    uint32_t* err = &d.cnt->err;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nthreads = gridDim.x * blockDim.x;
    const uint32_t n_man = d.cnt->manifolds;
    if (reference == 1u) {
        for (uint32_t m = tid; m < n_man; m += nthreads) ref_count(d, m);
        if (grid_sync(g, err)) return;
        if (blockIdx.x == 0) block_scan_range(d.adj_cnt, d.adj_start, d.nb);
        if (grid_sync(g, err)) return;
        for (uint32_t m = tid; m < n_man; m += nthreads) ref_scatter(d, m);
        if (grid_sync(g, err)) return;
        for (uint32_t x = tid; x < d.nb; x += nthreads) ref_chain(d, x);
        if (grid_sync(g, err)) return;
        for (uint32_t round = 0; round < max_rounds; ++round) {
            if (__hip_atomic_load(&d.cnt->remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) break;
            const uint32_t base = tid - (tid & 63u);
            for (uint32_t m0 = base; m0 < n_man; m0 += nthreads) {
                const uint32_t m = m0 + (tid & 63u);
                colour_taken(d, m < n_man ? ref_layer(d, m, err) : 0u);
            }
            if (grid_sync(g, err)) return;
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERR_COLOURS) break;
        }
        if (tid == 0 && __hip_atomic_load(&d.cnt->remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
            atomicOr(err, ERR_COLOURS);
        return;
    }
    const int prev = d.cur ^ 1;
    constexpr uint32_t kHinted = 0x80000000u;
    // A: claim last step's colour (mode 2: colour afresh, skipping A and B)
    for (uint32_t m = tid; reference == 0u && m < n_man; m += nthreads) {
        if (d.m_colour[m] != kUncoloured) continue;
        uint2 ab = d.m_ab[m];
        float4 cv;
        if (!cache_lookup(d, prev, pack_key(ab.x, ab.y, __float_as_uint(d.m_q[4 * m].y)), cv)) continue;
        uint32_t c = __float_as_uint(cv.w);
        if (c >= uint32_t(kMaskColours)) continue;
        unsigned long long bit = 1ull << c;
        d.m_colour[m] = kHinted | c;
        if (body_active(d.b_flags[ab.x]) && (atomicOr(&d.col_mask[ab.x], bit) & bit)) atomicOr(&d.col_dup[ab.x], bit);
        if (body_active(d.b_flags[ab.y]) && (atomicOr(&d.col_mask[ab.y], bit) & bit)) atomicOr(&d.col_dup[ab.y], bit);
    }
    if (grid_sync(g, err)) return;
    // B: keep the uncontested claims; a dropped claim clears its bit (no kept manifold shares it)
    if (reference == 0u) {
        uint32_t base = tid - (tid & 63u);
        for (uint32_t m0 = base; m0 < n_man; m0 += nthreads) {
            uint32_t m = m0 + (tid & 63u);
            uint32_t took = 0, top = 0;
            uint32_t col = m < n_man ? d.m_colour[m] : 0u;
            if (m < n_man && col != kUncoloured && (col & kHinted) && col < (kHinted | uint32_t(kMaskColours))) {
                uint32_t c = col & ~kHinted;
                unsigned long long bit = 1ull << c;
                uint2 ab = d.m_ab[m];
                bool aa = body_active(d.b_flags[ab.x]), ba = body_active(d.b_flags[ab.y]);
                bool lost = (aa && (d.col_dup[ab.x] & bit)) || (ba && (d.col_dup[ab.y] & bit));
                if (lost) {
                    d.m_colour[m] = kUncoloured;
                    if (aa) atomicAnd(&d.col_mask[ab.x], ~bit);
                    if (ba) atomicAnd(&d.col_mask[ab.y], ~bit);
                } else {
                    d.m_colour[m] = c;
                    took = 1u;
                    top = c + 1u;
                }
            }
            took = wave_sum(took);
            top = wave_max(top);
            if (lane0() && took) {
                atomicSub(&d.cnt->remaining, took);
                atomicMax(&d.cnt->colours, top);
            }
        }
    }
    if (grid_sync(g, err)) return;
    // C: Jones-Plassmann rounds for the rest
    const bool tiled = reference == 3u;
    for (uint32_t round = 1; round <= max_rounds; ++round) {
        if (__hip_atomic_load(&d.cnt->remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) break;
        for (uint32_t m = tid; m < n_man; m += nthreads) {
            if (d.m_colour[m] != kUncoloured) continue;
            uint2 ab = d.m_ab[m];
            const uint32_t po = tiled ? (d.m_tile[m] >> 8) * d.nb : 0u;
            unsigned long long key = colour_key(round, m, ab, d.m_shapes[m], d.m_meta[m].x & 255u);
            if (body_active(d.b_flags[ab.x])) atomicMax(&d.col_best[po + ab.x], key);
            if (body_active(d.b_flags[ab.y])) atomicMax(&d.col_best[po + ab.y], key);
        }
        if (grid_sync(g, err)) return;
        uint32_t base = tid - (tid & 63u);
        for (uint32_t m0 = base; m0 < n_man; m0 += nthreads) {
            uint32_t m = m0 + (tid & 63u);
            uint32_t took = 0, top = 0, ph = 0;
            if (m < n_man && d.m_colour[m] == kUncoloured) {
                uint2 ab = d.m_ab[m];
                ph = tiled ? d.m_tile[m] >> 8 : 0u;
                const uint32_t po = ph * d.nb;
                unsigned long long key = colour_key(round, m, ab, d.m_shapes[m], d.m_meta[m].x & 255u);
                bool aa = body_active(d.b_flags[ab.x]), ba = body_active(d.b_flags[ab.y]);
                bool mine = (!aa || d.col_best[po + ab.x] == key) && (!ba || d.col_best[po + ab.y] == key);
                if (mine) {
                    unsigned long long used = (aa ? d.col_mask[po + ab.x] : 0ull) | (ba ? d.col_mask[po + ab.y] : 0ull);
                    unsigned long long freeb = ~used;
                    if (freeb == 0ull) {
                        atomicOr(err, ERR_COLOURS);
                    } else {
                        uint32_t c = uint32_t(__ffsll(static_cast<long long>(freeb)) - 1);
                        d.m_colour[m] = c;
                        unsigned long long bit = 1ull << c;
                        if (aa) atomicOr(&d.col_mask[po + ab.x], bit);
                        if (ba) atomicOr(&d.col_mask[po + ab.y], bit);
                        took = 1u;
                        top = c + 1u;
                    }
                }
            }
            if (tiled) {
                for (uint32_t q = 0; q < 3u; ++q) {
                    const uint32_t tq = wave_max(ph == q ? top : 0u);
                    if (lane0() && tq) atomicMax(&d.cnt->tile_cols[q], tq);
                }
            }
            took = wave_sum(took);
            top = wave_max(top);
            if (lane0() && took) {
                atomicSub(&d.cnt->remaining, took);
                atomicMax(&d.cnt->colours, top);
            }
        }
        if (grid_sync(g, err)) return;
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERR_COLOURS) break;
    }
    // an uncoloured active manifold would be left out of the solve: report it
    if (tid == 0 && __hip_atomic_load(&d.cnt->remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
        atomicOr(err, ERR_COLOURS);
}

inline uint32_t blocks(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st) {
    if (!n) return;
    uint32_t g = uint32_t(std::min<size_t>((n + kBlock - 1) / kBlock, 4096));
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(kBlock), 0, st, p, v, n);
}
void launch_fill_u64(unsigned long long* p, unsigned long long v, size_t n, hipStream_t st) {
    if (!n) return;
    uint32_t g = uint32_t(std::min<size_t>((n + kBlock - 1) / kBlock, 4096));
    hipLaunchKernelGGL(k_fill_u64, dim3(g), dim3(kBlock), 0, st, p, v, n);
}
// exclusive scan of n elements (in == out allowed). scratch holds the block sums: ceil(n / 1024)
// words (>= 1), plus, when n > 2^20, what the scan of those block sums needs in turn.
void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, uint32_t* scratch, hipStream_t st) {
    uint32_t nb = (n + 1023) / 1024;
    if (nb == 0) nb = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, st, in, n, scratch);
    if (nb <= 1024) hipLaunchKernelGGL(k_scan_bsums, dim3(1), dim3(kBlock), 0, st, scratch, nb, total);
    else launch_scan(scratch, scratch, nb, total, scratch + nb, st);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kBlock), 0, st, in, out, n, scratch);
}
// scratch words launch_scan needs for n elements
static size_t scan_scratch_words(size_t n) {
    size_t nb = std::max<size_t>(1, (n + 1023) / 1024);
    return nb <= 1024 ? nb : nb + scan_scratch_words(nb);
}

void launch_init_bodies(const DevState& d, hipStream_t st) {
    if (d.nb) hipLaunchKernelGGL(k_init_bodies, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d);
}
void launch_scatter_bodies(const DevState& d, const uint32_t* idx, const float4* rec, const uint32_t* flags, uint32_t n,
                           int inertia, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_scatter_bodies, dim3(blocks(n)), dim3(kBlock), 0, st, d, idx, rec, flags, n, inertia);
}
void launch_init_proxies(const DevState& d, const StepParams& p, hipStream_t st) {
    if (d.ns) hipLaunchKernelGGL(k_init_proxies, dim3(blocks(d.ns)), dim3(kBlock), 0, st, d, p);
}
void launch_integrate_forces(const DevState& d, const StepParams& p, hipStream_t st) {
    hipLaunchKernelGGL(k_integrate_forces, dim3(blocks(d.nb ? d.nb : 1)), dim3(kBlock), 0, st, d, p);
}
// bounds, grid build, per-emitter pair counts and their scan (pair_off; total in cnt->pairs)
void launch_broadphase_count(const DevState& d, const StepParams& p, hipStream_t st) {
    if (!d.ns) return;
    hipLaunchKernelGGL(k_bounds, dim3(blocks(d.ns)), dim3(kBlock), 0, st, d, p);
    hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(1), 0, st, d);
    hipLaunchKernelGGL(k_grid_clear, dim3(256), dim3(kBlock), 0, st, d);
    hipLaunchKernelGGL(k_grid_count, dim3(blocks(d.ns)), dim3(kBlock), 0, st, d);
    launch_scan(d.cell_count, d.cell_start, kMaxCells, nullptr, d.scan_tmp, st);
    hipLaunchKernelGGL(k_grid_scatter, dim3(blocks(d.ns)), dim3(kBlock), 0, st, d);
    const uint32_t lanes = d.ns * uint32_t(kPairLanes);
    hipLaunchKernelGGL(k_pairs<0>, dim3(blocks(lanes)), dim3(kBlock), 0, st, d);
    launch_scan(d.pair_count, d.pair_off, lanes, &d.cnt->pairs, d.scan_tmp, st);
}
void launch_broadphase_write(const DevState& d, const StepParams& p, hipStream_t st) {
    (void)p;
    if (d.ns) {
        hipLaunchKernelGGL(k_pairs<1>, dim3(blocks(d.ns * uint32_t(kPairLanes))), dim3(kBlock), 0, st, d);
        hipLaunchKernelGGL(k_pairs_sort, dim3(blocks(d.ns)), dim3(kBlock), 0, st, d);
    }
}
void launch_narrowphase(const DevState& d, const StepParams& p, uint32_t n_pairs, hipStream_t st) {
    if (!n_pairs) return;
    hipLaunchKernelGGL(k_narrowphase, dim3(blocks(n_pairs)), dim3(kBlock), 0, st, d, p, n_pairs);
    if (d.n_meshes) hipLaunchKernelGGL(k_narrowphase_mesh, dim3(blocks(n_pairs)), dim3(kBlock), 0, st, d, n_pairs);
    launch_scan(d.pm_valid, d.pm_off, n_pairs, nullptr, d.scan_tmp, st);
    hipLaunchKernelGGL(k_compact, dim3(blocks(n_pairs)), dim3(kBlock), 0, st, d, n_pairs);
}
void launch_colour_init(const DevState& d, uint32_t n_man, hipStream_t st) {
    launch_fill_u64(d.col_best, 0ull, d.nb, st);
    launch_fill_u64(d.col_mask, 0ull, d.nb, st);
    launch_fill_u64(d.col_dup, 0ull, d.nb, st);
    launch_fill_u32(d.adj_cnt, 0u, d.nb, st);
    launch_fill_u32(d.adj_cur, 0u, d.nb, st);
    if (n_man) hipLaunchKernelGGL(k_colour_init, dim3(blocks(n_man)), dim3(kBlock), 0, st, d, n_man);
}
void launch_reference_chains(const DevState& d, uint32_t n_man, hipStream_t st) {
    if (!n_man || !d.nb) return;
    hipLaunchKernelGGL(k_ref_count, dim3(blocks(n_man)), dim3(kBlock), 0, st, d, n_man);
    launch_scan(d.adj_cnt, d.adj_start, d.nb, nullptr, d.scan_tmp, st);
    hipLaunchKernelGGL(k_ref_scatter, dim3(blocks(n_man)), dim3(kBlock), 0, st, d, n_man);
    hipLaunchKernelGGL(k_ref_chain, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d);
}
void launch_reference_layers(const DevState& d, uint32_t n_man, uint32_t rounds, hipStream_t st) {
    if (!n_man) return;
    for (uint32_t r = 0; r < rounds; ++r)
        hipLaunchKernelGGL(k_ref_layer, dim3(blocks(n_man)), dim3(kBlock), 0, st, d, n_man);
}
void launch_colour_rounds(const DevState& d, uint32_t n_man, uint32_t first_round, uint32_t rounds, hipStream_t st) {
    if (!n_man) return;
    for (uint32_t r = 0; r < rounds; ++r) {
        hipLaunchKernelGGL(k_colour_propose, dim3(blocks(n_man)), dim3(kBlock), 0, st, d, n_man, first_round + r);
        hipLaunchKernelGGL(k_colour_select, dim3(blocks(n_man)), dim3(kBlock), 0, st, d, n_man, first_round + r);
    }
}
size_t col_table_words(size_t max_blocks) {
    const size_t n = size_t(kSortBuckets) * max_blocks;
    return n + scan_scratch_words(n) + 1;
}
void launch_build_order(const DevState& d, uint32_t n_man, hipStream_t st) {
    uint32_t nblk = blocks(n_man ? n_man : 1);
    if (n_man) hipLaunchKernelGGL(k_colour_hist, dim3(nblk), dim3(kBlock), 0, st, d, n_man, nblk);
    else launch_fill_u32(d.col_table, 0u, size_t(kSortBuckets) * nblk, st);
    // col_table_off doubles as scan scratch elsewhere: scan into it, block sums in col_table's tail
    // (sized by col_table_words)
    uint32_t n = kSortBuckets * nblk;
    uint32_t* scratch = d.col_table + n;
    uint32_t* total = scratch + scan_scratch_words(n);
    launch_scan(d.col_table, d.col_table_off, n, total, scratch, st);
    if (n_man) hipLaunchKernelGGL(k_colour_scatter, dim3(nblk), dim3(kBlock), 0, st, d, n_man, nblk);
    hipLaunchKernelGGL(k_colour_offsets, dim3(1), dim3(kBlock), 0, st, d, nblk, total);
}
void launch_ready_split(const DevState& d, uint32_t bound, hipStream_t st) {
    if (!bound) return;
    if (d.nb) launch_fill_u32(d.body_last, 0u, d.nb, st);
    launch_fill_u32(d.ready_cnt, 0u, 2, st);
    hipLaunchKernelGGL(k_ready_last, dim3(blocks(bound)), dim3(kBlock), 0, st, d, bound);
    hipLaunchKernelGGL(k_ready_split, dim3(blocks(bound)), dim3(kBlock), 0, st, d, bound);
}
void launch_solver_setup(const DevState& d, const StepParams& p, uint32_t bound, hipStream_t st) {
    // solver bodies first: the point rows are computed from them
    if (d.nb) hipLaunchKernelGGL(k_solver_bodies, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d);
    if (bound) hipLaunchKernelGGL(k_solver_setup, dim3(blocks(bound)), dim3(kBlock), 0, st, d, p, bound);
}
void launch_apply_predictions(const DevState& d, uint32_t n_bound, hipStream_t st) {
    if (!n_bound || !d.mlp) return;
    launch_mlp_infer_dev(d.mlp, d.feat, d.pred, &d.cnt->predicted, n_bound, st);
    hipLaunchKernelGGL(k_apply_predictions, dim3(blocks(n_bound)), dim3(kBlock), 0, st, d, n_bound);
}
void launch_warm_apply(const DevState& d, uint32_t begin, uint32_t end, hipStream_t st) {
    if (end > begin) hipLaunchKernelGGL(k_warm_apply, dim3(blocks(end - begin)), dim3(kBlock), 0, st, d, begin, end);
}
void launch_velocity(const DevState& d, const StepParams& p, uint32_t begin, uint32_t end, int use_rest, uint32_t offset,
                     hipStream_t st) {
    if (end > begin)
        hipLaunchKernelGGL(k_velocity, dim3(blocks(end - begin)), dim3(kBlock), 0, st, d, p, begin, end, use_rest, offset);
}
void launch_residual(const DevState& d, const StepParams& p, uint32_t n_active, int use_rest, hipStream_t st) {
    hipLaunchKernelGGL(k_residual_reset, dim3(1), dim3(1), 0, st, d);
    if (n_active) hipLaunchKernelGGL(k_residual, dim3(blocks(n_active)), dim3(kBlock), 0, st, d, p, n_active, use_rest);
}
void launch_set_passes(const DevState& d, uint32_t passes, hipStream_t st) {
    hipLaunchKernelGGL(k_set_passes, dim3(1), dim3(1), 0, st, d, passes);
}
void launch_position_setup(const DevState& d, hipStream_t st) {
    if (d.nb) hipLaunchKernelGGL(k_position_setup, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d);
}
void launch_position(const DevState& d, const StepParams& p, uint32_t begin, uint32_t end, hipStream_t st) {
    if (end > begin) hipLaunchKernelGGL(k_position, dim3(blocks(end - begin)), dim3(kBlock), 0, st, d, p, begin, end);
}
void launch_solver_finish(const DevState& d, uint32_t bound, hipStream_t st) {
    if (d.nb) hipLaunchKernelGGL(k_bodies_back, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d);
    size_t cap = size_t(d.cache_mask) + 1;
    launch_fill_u64(d.ck_keys[d.cur], kEmptyKey, cap, st);
    launch_fill_u64(d.ck_ord[d.cur], 0ull, cap, st);
    if (bound) {
        hipLaunchKernelGGL(k_cache_insert, dim3(blocks(bound)), dim3(kBlock), 0, st, d, bound);
        hipLaunchKernelGGL(k_cache_write, dim3(blocks(bound)), dim3(kBlock), 0, st, d, bound);
    }
}
void launch_islands_sleep(const DevState& d, const StepParams& p, uint32_t bound, hipStream_t st) {
    if (!d.nb) return;
    hipLaunchKernelGGL(k_uf_init, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d);
    if (bound) hipLaunchKernelGGL(k_uf_union, dim3(blocks(bound)), dim3(kBlock), 0, st, d, bound);
    if (d.nj) hipLaunchKernelGGL(k_uf_union_joints, dim3(blocks(d.nj)), dim3(kBlock), 0, st, d);
    hipLaunchKernelGGL(k_uf_roots, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d);
    hipLaunchKernelGGL(k_sleep_timers, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d, p);
    if (bound) hipLaunchKernelGGL(k_kin_wake, dim3(blocks(bound)), dim3(kBlock), 0, st, d, p, bound);
    if (d.nj) hipLaunchKernelGGL(k_kin_wake_joints, dim3(blocks(d.nj)), dim3(kBlock), 0, st, d, p);
    hipLaunchKernelGGL(k_island_min, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d);
    hipLaunchKernelGGL(k_sleep, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d, p);
}
// cooperative launches; returns the HIP error of the launch. The barrier count is zeroed by a
// kernel (not a memset) so the launch sequence may be captured in a graph.
hipError_t launch_solve_persistent(const DevState& d, const StepParams& p, const uint32_t cfgv[8], float tol,
                                   uint32_t nblocks, hipStream_t st) {
    SolveCfg c;
    c.vel_iters = cfgv[0]; c.relax_iters = cfgv[1]; c.pos_iters = cfgv[2];
    c.residual_mode = cfgv[3]; c.residual_max = cfgv[4]; c.residual_tol = tol;
    c.one_xcd = cfgv[5];
    c.ready = cfgv[6];
    c.check_every = cfgv[7] ? cfgv[7] : 1u;
    c.tiled_lds = 0u;
    launch_fill_u32(d.bar, 0u, kBarWords, st);
    DevState dd = d;
    StepParams pp = p;
    uint32_t* bar = d.bar;
    void* args[] = {&dd, &pp, &c, &bar};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_solve_persistent), dim3(nblocks),
                                      dim3(kPersistBlock), args, 0, st);
}
hipError_t launch_solve_tiled(const DevState& d, const StepParams& p, const uint32_t cfgv[8], float tol, uint32_t nblocks,
                              hipStream_t st) {
    static bool attr = false;  // the dynamic LDS beyond the default 64 KB, set once per process
    if (!attr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_solve_tiled),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(kTileLds));
        if (e != hipSuccess) return e;
        attr = true;
    }
    SolveCfg c;
    c.vel_iters = cfgv[0]; c.relax_iters = cfgv[1]; c.pos_iters = cfgv[2];
    c.residual_mode = cfgv[3]; c.residual_max = cfgv[4]; c.residual_tol = tol;
    c.one_xcd = cfgv[5];
    c.ready = 0u;
    c.check_every = cfgv[7] ? cfgv[7] : 1u;
    c.tiled_lds = cfgv[6];
    launch_fill_u32(d.bar, 0u, kBarWords, st);
    DevState dd = d;
    StepParams pp = p;
    uint32_t* bar = d.bar;
    void* args[] = {&dd, &pp, &c, &bar};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_solve_tiled), dim3(nblocks), dim3(kTileBlock),
                                      args, kTileLds, st);
}
void launch_solve_bytes(const DevState& d, const uint32_t cfgv[8], uint32_t bound, hipStream_t st) {
    if (bound) hipLaunchKernelGGL(k_solve_bytes, dim3(blocks(bound)), dim3(kBlock), 0, st, d, cfgv[0], cfgv[1], cfgv[2], cfgv[3]);
}
hipError_t launch_colour_persistent(const DevState& d, uint32_t max_rounds, uint32_t nblocks, uint32_t reference,
                                    hipStream_t st) {
    launch_fill_u32(d.bar, 0u, 1, st);
    DevState dd = d;
    uint32_t* bar = d.bar;
    void* args[] = {&dd, &max_rounds, &reference, &bar};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_colour_persistent), dim3(nblocks),
                                      dim3(kPersistBlock), args, 0, st);
}
int max_coresident_blocks(int device) {
    int per_cu = 0, per_cu2 = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_solve_persistent),
                                                     kPersistBlock, 0) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, reinterpret_cast<const void*>(&k_colour_persistent),
                                                     kPersistBlock, 0) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    return std::min(per_cu, per_cu2) * cus;
}

void launch_integrate_velocities(const DevState& d, const StepParams& p, hipStream_t st) {
    if (d.nb) hipLaunchKernelGGL(k_integrate_velocities, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d, p);
}
void launch_wake_jointed(const DevState& d, hipStream_t st) {
    if (d.nj) hipLaunchKernelGGL(k_wake_jointed, dim3(1), dim3(1024), 0, st, d);
}
void launch_joint_warm(const DevState& d, uint32_t begin, uint32_t end, hipStream_t st) {
    if (end > begin) hipLaunchKernelGGL(k_joint_warm, dim3(blocks(end - begin)), dim3(kBlock), 0, st, d, begin, end);
}
void launch_joint_velocity(const DevState& d, const StepParams& p, int rev, uint32_t begin, uint32_t end, hipStream_t st) {
    if (end > begin)
        hipLaunchKernelGGL(k_joint_velocity, dim3(blocks(end - begin)), dim3(kBlock), 0, st, d, p, rev, begin, end);
}
void launch_joint_position(const DevState& d, const StepParams& p, uint32_t begin, uint32_t end, hipStream_t st) {
    if (end > begin) hipLaunchKernelGGL(k_joint_position, dim3(blocks(end - begin)), dim3(kBlock), 0, st, d, p, begin, end);
}
void launch_joints_post(const DevState& d, const StepParams& p, uint32_t pos_iters, hipStream_t st) {
    if (!d.nj) return;
    hipLaunchKernelGGL(k_joints_post, dim3(1), dim3(1024), 0, st, d, p, pos_iters > 1u ? pos_iters : 1u);
    if (d.nb) hipLaunchKernelGGL(k_update_inertia, dim3(blocks(d.nb)), dim3(kBlock), 0, st, d);
}

}  // namespace pxg
