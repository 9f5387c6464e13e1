// Tiled solve order (PXG_ORDER_TILED) for large piles: set-up kernels of one step.
//
// The awake dynamic bodies are split into kTiles spatial tiles of equal count (kTileNx strips by x
// rank, each cut into kTileNy tiles by y rank). A manifold whose awake dynamic bodies lie in one
// tile is in phase 0 (group = that tile). The others cross a tile boundary: their bodies are joined
// into connected components (union-find over the crossing manifolds), and each component goes
// whole to one of kTiles groups of phase 1 (components dealt by size, in snake order). Groups of one phase share no
// awake dynamic body, so each group's colours can run in one workgroup with workgroup barriers
// only, all groups of a phase at once. (Phase 2, one group, is kept by the solve for generality;
// this set-up leaves it empty.)
//
// The colouring (Jones-Plassmann, per-phase masks) gives each manifold a colour within its phase;
// the solve order is sorted by (global colour = colours of the earlier phases + colour, group,
// more contact points first, manifold), so it is colour-major and every colour is a valid colour
// of the whole world (its manifolds share no awake dynamic body): the sequential solve of that
// order is what the tiled solve computes. Within a (colour, group) range the manifolds with more
// points come first, so the solve can give them more lanes each. t_off gives every (global colour, group) range; per (phase, group) a table of
// every body its manifolds touch (tb_list, ascending; a body that is not awake dynamic, such as the
// ground, can be in several groups' tables), and each solve position's two indices in its group's
// table (sm_loc), let the solve stage a group's bodies in LDS.
//
// Every step: 5 radix sorts (bodies by x; bodies by (strip, y); crossing components by size;
// manifolds by (colour, group, points); manifold sides by (phase, group, body)), each over at most
// max(bodies, 2 pairs) 32/64-bit keys, one union-find over the crossing manifolds and one scan (the distinct
// sides).
#include <rocprim/device/device_radix_sort.hpp>

#include "pxg_state.h"

namespace pxg {

namespace {

constexpr uint32_t kColUncoloured = 0xFFFFFFFFu, kColSkipped = 0xFFFFFFFEu;  // the colouring's markers
constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ bool awake_dynamic(uint32_t f) {
    return (f & BF_TYPE) == 2u && (f & BF_AWAKE) && (f & BF_ALIVE);
}
// order-preserving float -> uint
__device__ __forceinline__ uint32_t ord_bits(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
// first index in sorted keys[0, n) whose key is >= v
template <class K>
__device__ __forceinline__ uint32_t lower_bound(const K* keys, uint32_t n, K v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// sort keys of the bodies by x (awake dynamic bodies first, by x; the rest last); counts them
__global__ void k_tile_xkeys(DevState d) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    bool act = false;
    if (b < d.nb) {
        act = awake_dynamic(d.b_flags[b]);
        const uint32_t k = ord_bits(d.b_pos[b].x);
        d.tk_a[b] = act ? (k < kNone ? k : kNone - 1u) : kNone;
        d.tv_a[b] = b;
    }
    uint32_t c = act ? 1u : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63u) == 0u && c) atomicAdd(&d.cnt->tile_n, c);
}
// from the x rank r of each awake dynamic body: its strip, and the (strip, y) keys of the second
// sort
__global__ void k_tile_strips(DevState d) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d.nb) return;
    const uint32_t n = d.cnt->tile_n;
    const uint32_t b = d.tv_b[r];
    unsigned long long k = ~0ull;
    if (r < n) {
        const unsigned long long s = (static_cast<unsigned long long>(r) * kTileNx) / n;
        k = (s << 32) | ord_bits(d.b_pos[b].y);
    }
    d.tk64_a[r] = k;
    d.tv_a[r] = b;
}
// y rank within the strip -> tile of each awake dynamic body
__global__ void k_tile_ranks(DevState d) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = d.cnt->tile_n;
    if (p >= n) return;
    const unsigned long long strip = d.tk64_b[p] >> 32;
    const uint32_t lo = lower_bound(d.tk64_b, n, strip << 32);
    const uint32_t hi = lower_bound(d.tk64_b, n, (strip + 1ull) << 32);
    const uint32_t rank = p - lo, cnt = hi - lo;
    const uint32_t ty = uint32_t((static_cast<unsigned long long>(rank) * kTileNy) / cnt);
    d.t_body[d.tv_b[p]] = uint32_t(strip) * kTileNy + ty;
}
// union-find over the manifolds that cross a tile boundary (the larger root is hooked under the
// smaller, so every component's root is its smallest body whatever the order of the unions)
__device__ __forceinline__ uint32_t tile_find(uint32_t* parent, uint32_t x) {
    while (true) {
        const uint32_t p = parent[x];
        if (p == x) return x;
        const uint32_t g = parent[p];
        if (g != p) parent[x] = g;  // path halving (a benign race: g is an ancestor of x)
        x = p;
    }
}
__global__ void k_tile_uf_init(DevState d) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < d.nb) d.t_par[b] = b;
}
__global__ void k_tile_cross(DevState d, uint32_t n_bound) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_bound || m >= d.cnt->manifolds || d.m_colour[m] != kColUncoloured) return;
    const uint2 ab = d.m_ab[m];
    if (!awake_dynamic(d.b_flags[ab.x]) || !awake_dynamic(d.b_flags[ab.y])) return;
    if (d.t_body[ab.x] == d.t_body[ab.y]) return;
    uint32_t a = tile_find(d.t_par, ab.x), b = tile_find(d.t_par, ab.y);
    for (int guard = 0; guard < (1 << 24) && a != b; ++guard) {
        if (a > b) { const uint32_t t = a; a = b; b = t; }
        const uint32_t old = atomicCAS(&d.t_par[b], b, a);
        if (old == b) return;
        b = tile_find(d.t_par, old);
        a = tile_find(d.t_par, a);
    }
}
// each body's component root (read-only over the parent array), into t_body's second half
__global__ void k_tile_uf_roots(DevState d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nb) return;
    uint32_t x = i;
    for (uint32_t p = d.t_par[x]; p != x; p = d.t_par[x]) x = p;
    d.t_body[d.nb + i] = x;
}
// crossing manifolds per component (counted at the root)
__global__ void k_tile_csize(DevState d, uint32_t n_bound) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_bound || m >= d.cnt->manifolds || d.m_colour[m] != kColUncoloured) return;
    const uint2 ab = d.m_ab[m];
    if (!awake_dynamic(d.b_flags[ab.x]) || !awake_dynamic(d.b_flags[ab.y])) return;
    if (d.t_body[ab.x] == d.t_body[ab.y]) return;
    atomicAdd(&d.t_csz[d.t_body[d.nb + ab.x]], 1u);
}
// sort keys of the components: larger first, then by root; other bodies last
__global__ void k_tile_ckeys(DevState d) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.nb) return;
    const uint32_t c = d.t_csz[b];
    d.tk64_a[b] = c ? ((static_cast<unsigned long long>(~c) << 32) | b) : ~0ull;
    d.tv_a[b] = b;
    if (c) atomicAdd(&d.cnt->tile_ncomp, 1u);
}
// the components dealt to the groups in snake order of size (0, 1, .., 31, 31, 30, .., 0, 0, 1, ..),
// which keeps the groups' manifold counts close
__global__ void k_tile_cgroup(DevState d) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.cnt->tile_ncomp) return;
    const uint32_t round = k / kTiles, pos = k % kTiles;
    d.t_csz[d.tv_b[k]] = (round & 1u) ? kTiles - 1u - pos : pos;
}
// phase and group of each active (still uncoloured) manifold, and the table keys of its two bodies
// (phase * kTiles + group) << 20 | body (none for the other manifolds)
__global__ void k_tile_phase(DevState d, uint32_t n_bound) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_bound) return;
    uint32_t ka = kNone, kb = kNone;
    if (m < d.cnt->manifolds && d.m_colour[m] == kColUncoloured) {
        const uint2 ab = d.m_ab[m];
        const bool aa = awake_dynamic(d.b_flags[ab.x]), ba = awake_dynamic(d.b_flags[ab.y]);
        const uint32_t nb = d.nb;
        const uint32_t ta = aa ? d.t_body[ab.x] : kNone, tb = ba ? d.t_body[ab.y] : kNone;
        uint32_t ph, g;
        if (!aa || !ba || ta == tb) {
            ph = 0u;
            g = aa ? ta : tb;
        } else {
            // crossing: its component's group (both bodies have the same root)
            ph = 1u;
            g = d.t_csz[d.t_body[nb + ab.x]];
        }
        d.m_tile[m] = (ph << 8) | g;
        const uint32_t grp = (ph * kTiles + g) << 20;
        ka = grp | ab.x;
        kb = grp | ab.y;
    }
    d.ts_a[2u * m] = ka;
    d.ts_a[2u * m + 1u] = kb;
}
// solve-order keys: (global colour * kTiles + group) << 20 | manifold; the skipped group after
// every colour; triggers, uncoloured and unused entries last
// order keys: bucket (global colour * kTiles + group; kTileSkipBucket for skipped manifolds, one
// more for none) << 2 | (4 - points); the value is the manifold, and the stable sort keeps
// manifold order within a key
constexpr uint32_t kOrderKeyBits = 14u;
__global__ void k_tile_keys(DevState d, uint32_t n_bound) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_bound) return;
    uint32_t key = (kTileSkipBucket + 1u) << 2;
    if (m < d.cnt->manifolds) {
        const uint32_t c = d.m_colour[m];
        const uint32_t cnt = d.m_meta[m].x & 255u;
        const uint32_t fewer = cnt >= 1u && cnt <= 4u ? 4u - cnt : 3u;
        if (c == kColSkipped) {
            key = (kTileSkipBucket << 2) | fewer;
        } else if (c < uint32_t(kMaskColours)) {
            const uint32_t t = d.m_tile[m], ph = t >> 8, g = t & 255u;
            const uint32_t* cols = d.cnt->tile_cols;
            const uint32_t base = ph == 0u ? 0u : (ph == 1u ? cols[0] : cols[0] + cols[1]);
            const uint32_t gc = base + c;
            if (gc >= kTileColours) atomicOr(&d.cnt->err, ERR_COLOURS);
            else key = ((gc * kTiles + g) << 2) | fewer;
        }
    }
    d.tk_a[m] = key;
    d.tv_a[m] = m;
}
// bucket offsets (t_off), the colour table, the order and the solver counts
__global__ void k_tile_offsets(DevState d, uint32_t n_bound) {
    for (uint32_t b = threadIdx.x; b <= kTileSkipBucket + 1u; b += blockDim.x)
        d.t_off[b] = lower_bound(d.tk_b, n_bound, b << 2);
}
__global__ void k_tile_order(DevState d, uint32_t n_bound) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t active = d.t_off[kTileSkipBucket], ordered = d.t_off[kTileSkipBucket + 1u];
    if (i < ordered && i < n_bound) d.order[i] = d.tv_b[i];
    if (i <= uint32_t(kBuckets)) {
        const uint32_t c = i;
        d.cnt->colour_off[c] = c == uint32_t(kBuckets) ? ordered
                               : (c <= kTileColours ? d.t_off[c * kTiles] : active);
    }
    if (i == 0u) {
        const uint32_t* cols = d.cnt->tile_cols;
        d.cnt->active = active;
        d.cnt->ordered = ordered;
        d.cnt->colours = cols[0] + cols[1] + cols[2];
    }
}
// the distinct sorted side keys (the bodies of each group's table): first occurrences
__global__ void k_tile_uflags(DevState d, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = d.tk_a[i];
    d.tf[i] = (k != kNone && (i == 0u || d.tk_a[i - 1u] != k)) ? 1u : 0u;
}
__global__ void k_tile_compact(DevState d, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !d.tf[i]) return;
    const uint32_t k = d.tk_a[i], u = d.tu[i];
    d.tb_key[u] = k;
    d.tb_list[u] = k & 0xFFFFFu;
}
__global__ void k_tile_boffsets(DevState d) {
    const uint32_t nu = d.cnt->tile_nu;
    for (uint32_t k = threadIdx.x; k <= 3u * kTiles; k += blockDim.x) d.tb_off[k] = lower_bound(d.tb_key, nu, k << 20);
}
// per solve position: both bodies' indices in the group's table (0xFFFF past 16 bits, which only
// a table too large for a workgroup's LDS has)
__global__ void k_tile_slots(DevState d, uint32_t n_bound) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_bound || s >= d.cnt->active) return;
    const uint32_t m = d.order[s];
    const uint32_t t = d.m_tile[m], grp = (t >> 8) * kTiles + (t & 255u);
    const uint32_t lo = d.tb_off[grp], n = d.tb_off[grp + 1u] - lo;
    const uint2 ab = d.m_ab[m];
    const uint32_t la = lower_bound(d.tb_key + lo, n, (grp << 20) | ab.x);
    const uint32_t lb = lower_bound(d.tb_key + lo, n, (grp << 20) | ab.y);
    d.sm_loc[s] = (la < 0xFFFFu ? la : 0xFFFFu) | ((lb < 0xFFFFu ? lb : 0xFFFFu) << 16);
}
__global__ void k_tile_reset(DevState d) {
    if (threadIdx.x == 0) {
        d.cnt->tile_n = 0u;
        d.cnt->tile_ncomp = 0u;
        d.cnt->tile_wg = 0u;
        d.cnt->tile_need = 0u;
        d.cnt->tile_cols[0] = d.cnt->tile_cols[1] = d.cnt->tile_cols[2] = 0u;
    }
}
__global__ void k_tile_fill(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}
__global__ void k_tile_fill64(unsigned long long* p, size_t n) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0ull;
}

inline uint32_t nblocks(size_t n) { return uint32_t(std::min<size_t>((n + kBlock - 1) / kBlock, 65535)); }

}  // namespace

size_t tile_sort_bytes(uint32_t cap_bodies, uint32_t cap_items) {
    size_t a = 0, b = 0, c = 0, e = 0;
    uint32_t* k32 = nullptr;
    unsigned long long* k64 = nullptr;
    if (rocprim::radix_sort_pairs(nullptr, a, k32, k32, k32, k32, cap_bodies, 0u, 32u) != hipSuccess) return 0;
    if (rocprim::radix_sort_pairs(nullptr, b, k64, k64, k32, k32, cap_bodies, 0u, 64u) != hipSuccess) return 0;
    if (rocprim::radix_sort_keys(nullptr, c, k32, k32, cap_items, 0u, 32u) != hipSuccess) return 0;
    if (rocprim::radix_sort_pairs(nullptr, e, k32, k32, k32, k32, cap_items, 0u, kOrderKeyBits) != hipSuccess) return 0;
    return std::max(std::max(a, e), std::max(b, c));
}

hipError_t launch_tile_setup(const DevState& d, uint32_t n_bound, hipStream_t st) {
    const uint32_t nb = d.nb;
    hipLaunchKernelGGL(k_tile_reset, dim3(1), dim3(64), 0, st, d);
    hipLaunchKernelGGL(k_tile_fill, dim3(nblocks(2u * size_t(nb) + 1u)), dim3(kBlock), 0, st, d.t_body, kNone,
                       2u * size_t(nb));
    hipLaunchKernelGGL(k_tile_fill64, dim3(nblocks(3u * size_t(nb) + 1u)), dim3(kBlock), 0, st, d.t_col_best,
                       3u * size_t(nb));
    hipLaunchKernelGGL(k_tile_fill64, dim3(nblocks(3u * size_t(nb) + 1u)), dim3(kBlock), 0, st, d.t_col_mask,
                       3u * size_t(nb));
    if (!nb) return hipSuccess;
    hipLaunchKernelGGL(k_tile_xkeys, dim3(nblocks(nb)), dim3(kBlock), 0, st, d);
    size_t bytes = d.t_sort_bytes;
    hipError_t e = rocprim::radix_sort_pairs(d.t_sort_tmp, bytes, d.tk_a, d.tk_b, d.tv_a, d.tv_b, nb, 0u, 32u, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tile_strips, dim3(nblocks(nb)), dim3(kBlock), 0, st, d);
    bytes = d.t_sort_bytes;
    e = rocprim::radix_sort_pairs(d.t_sort_tmp, bytes, d.tk64_a, d.tk64_b, d.tv_a, d.tv_b, nb, 0u, 35u, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tile_ranks, dim3(nblocks(nb)), dim3(kBlock), 0, st, d);
    hipLaunchKernelGGL(k_tile_uf_init, dim3(nblocks(nb)), dim3(kBlock), 0, st, d);
    if (n_bound) hipLaunchKernelGGL(k_tile_cross, dim3(nblocks(n_bound)), dim3(kBlock), 0, st, d, n_bound);
    hipLaunchKernelGGL(k_tile_uf_roots, dim3(nblocks(nb)), dim3(kBlock), 0, st, d);
    // components sized, sorted by size, dealt to the groups (t_csz then holds each root's group)
    hipLaunchKernelGGL(k_tile_fill, dim3(nblocks(nb)), dim3(kBlock), 0, st, d.t_csz, 0u, size_t(nb));
    if (n_bound) hipLaunchKernelGGL(k_tile_csize, dim3(nblocks(n_bound)), dim3(kBlock), 0, st, d, n_bound);
    hipLaunchKernelGGL(k_tile_ckeys, dim3(nblocks(nb)), dim3(kBlock), 0, st, d);
    bytes = d.t_sort_bytes;
    e = rocprim::radix_sort_pairs(d.t_sort_tmp, bytes, d.tk64_a, d.tk64_b, d.tv_a, d.tv_b, nb, 0u, 64u, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tile_cgroup, dim3(nblocks(nb)), dim3(kBlock), 0, st, d);
    if (n_bound) hipLaunchKernelGGL(k_tile_phase, dim3(nblocks(n_bound)), dim3(kBlock), 0, st, d, n_bound);
    return hipSuccess;
}

hipError_t launch_tile_order(const DevState& d, uint32_t n_bound, hipStream_t st) {
    const uint32_t nb = d.nb;
    const uint32_t n = n_bound ? n_bound : 1u;
    hipLaunchKernelGGL(k_tile_keys, dim3(nblocks(n)), dim3(kBlock), 0, st, d, n);
    size_t bytes = d.t_sort_bytes;
    hipError_t e = rocprim::radix_sort_pairs(d.t_sort_tmp, bytes, d.tk_a, d.tk_b, d.tv_a, d.tv_b, n, 0u, kOrderKeyBits, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tile_offsets, dim3(1), dim3(1024), 0, st, d, n);
    hipLaunchKernelGGL(k_tile_order, dim3(nblocks(std::max<uint32_t>(n, kBuckets + 1))), dim3(kBlock), 0, st, d, n);
    if (!nb) return hipSuccess;
    // the group tables: sides sorted, duplicates dropped
    const uint32_t n2 = 2u * n;
    bytes = d.t_sort_bytes;
    e = rocprim::radix_sort_keys(d.t_sort_tmp, bytes, d.ts_a, d.tk_a, n2, 0u, 32u, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tile_uflags, dim3(nblocks(n2)), dim3(kBlock), 0, st, d, n2);
    launch_scan(d.tf, d.tu, n2, &d.cnt->tile_nu, d.scan_tmp, st);
    hipLaunchKernelGGL(k_tile_compact, dim3(nblocks(n2)), dim3(kBlock), 0, st, d, n2);
    hipLaunchKernelGGL(k_tile_boffsets, dim3(1), dim3(256), 0, st, d);
    return hipSuccess;
}

void launch_tile_slots(const DevState& d, uint32_t n_bound, hipStream_t st) {
    if (n_bound) hipLaunchKernelGGL(k_tile_slots, dim3(nblocks(n_bound)), dim3(kBlock), 0, st, d, n_bound);
}

}  // namespace pxg
