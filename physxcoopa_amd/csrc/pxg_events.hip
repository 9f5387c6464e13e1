// Contact events of the single-world step: emit_contact_events_ (world.h:1998-2018) with
// PairCache::classify (broadphase/pair_cache.h:63-82) on the GPU.
//
// Each step keeps a snapshot of its confirmed shape pairs -- the pairs with a valid manifold --
// as 64-bit keys kind << 34 | shape a << 17 | shape b (kind 0 solid, 1 trigger), sorted by a
// radix sort. Sorted this way the snapshot is the solid pair set in (a, b) order followed by the
// trigger pair set in (a, b) order, the two caches of the reference (world.h:2112-2113), and a
// pair of one kind is found in the other step's snapshot by one binary search over the full key.
// An entry of this step's snapshot is a stay when last step's holds the same key, else an enter;
// an entry of last step's snapshot missing from this step's is an exit. The queue order is the
// reference's: the solid cache, then the trigger cache, each as all enters, all stays, all exits,
// every group in pair order. That order is category-major, so one exclusive scan over the per-
// (category, block) counts gives every event its queue position.
#include "pxg_state.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace pxg {
namespace {

constexpr uint32_t kEvNone = 0xFFu;
constexpr uint32_t kEvBlock = 1024u;  // snapshot entries per block: 4 per thread
constexpr uint32_t kShapeBits = 17u;
constexpr unsigned long long kShapeMask = (1ull << kShapeBits) - 1ull;
constexpr unsigned kKeyBits = 2u * kShapeBits + 1u;
constexpr unsigned long long kNoKey = (1ull << kKeyBits) - 1ull;  // sorts last; no shape has slot 2^17 - 1
static_assert(kMaxShapes < kShapeMask, "shape slots must fit the key's 17-bit fields");

__device__ __forceinline__ uint32_t key_a(unsigned long long k) { return uint32_t((k >> kShapeBits) & kShapeMask); }
__device__ __forceinline__ uint32_t key_b(unsigned long long k) { return uint32_t(k & kShapeMask); }
__device__ __forceinline__ uint32_t key_kind(unsigned long long k) { return uint32_t(k >> (2u * kShapeBits)); }

__device__ __forceinline__ bool contains(const unsigned long long* list, uint32_t n, unsigned long long key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (list[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && list[lo] == key;
}

// emit_contact_events_'s push filter (world.h:2001-2007): both shapes in the table, both alive
// for an enter or a stay, both owners alive
__device__ __forceinline__ bool event_allowed(const DevState& d, uint32_t a, uint32_t b, bool is_exit) {
    if (a >= d.ns || b >= d.ns) return false;
    uint4 ia = d.s_info[a], ib = d.s_info[b];
    if (!is_exit && !((ia.y & SF_ENABLED) && (ib.y & SF_ENABLED))) return false;
    if (ia.x >= d.nb || ib.x >= d.nb) return false;
    return (d.b_flags[ia.x] & BF_ALIVE) && (d.b_flags[ib.x] & BF_ALIVE);
}

// Categories: 0 solid enter, 1 solid stay, 2 solid exit, 3 trigger enter, 4 trigger stay,
// 5 trigger exit. A thread's counts are packed in four 16-bit fields (a block has at most 1024
// events of one category): blocks over this step's snapshot count 0, 1, 3, 4; blocks over last
// step's count 2 and 5.
__device__ __forceinline__ uint32_t ev_field(uint32_t c) { return c == 2u ? 0u : c == 5u ? 1u : (c < 2u ? c : c - 1u); }

template <class T>
__device__ T block_exclusive_scan(T v, T* total) {
    __shared__ T wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wid) off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    if (total) *total = tot;
    return off + x - v;
}

// The unsorted keys: one per compacted manifold (pair order), kNoKey for the rest of the bound.
__global__ void __launch_bounds__(kBlock) k_ev_keys(DevState d, uint32_t bound, int cur) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t m = d.cnt->manifolds < d.cap_manifolds ? d.cnt->manifolds : d.cap_manifolds;
    if (k == 0) d.cnt->ev_snap_n[cur] = m < bound ? m : bound;
    if (k >= bound) return;
    unsigned long long key = kNoKey;
    if (k < m) {
        uint2 s = d.m_shapes[k];
        unsigned long long kind = (d.m_meta[k].x >> 8) & 1u;
        key = (kind << (2u * kShapeBits)) | (static_cast<unsigned long long>(s.x) << kShapeBits) | s.y;
    }
    d.ev_keys[k] = key;
}

__global__ void k_ev_empty(DevState d, int cur) { d.cnt->ev_snap_n[cur] = 0u; }

// Blocks [0, nbc) cover this step's snapshot, blocks [nbc, nblk) last step's.
__global__ void __launch_bounds__(kBlock) k_ev_classify(DevState d, uint32_t nbc, uint32_t nblk, int cur) {
    const bool mine = blockIdx.x < nbc;
    const uint32_t base = (mine ? blockIdx.x : blockIdx.x - nbc) * kEvBlock + threadIdx.x * 4u;
    const uint32_t n_cur = d.cnt->ev_snap_n[cur], n_prev = d.cnt->ev_snap_n[cur ^ 1];
    const unsigned long long* snap_cur = d.ev_snap[cur];
    const unsigned long long* snap_prev = d.ev_snap[cur ^ 1];
    unsigned long long packed = 0;
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t i = base + k;
        uint32_t c = kEvNone;
        if (mine) {
            if (i >= n_cur) break;
            const unsigned long long key = snap_cur[i];
            if (event_allowed(d, key_a(key), key_b(key), false))
                c = 3u * key_kind(key) + (contains(snap_prev, n_prev, key) ? 1u : 0u);
            d.ev_code[i] = uint8_t(c);
        } else {
            if (i >= n_prev) break;
            const unsigned long long key = snap_prev[i];
            if (!contains(snap_cur, n_cur, key) && event_allowed(d, key_a(key), key_b(key), true))
                c = 3u * key_kind(key) + 2u;
            d.ev_code[d.cap_pairs + i] = uint8_t(c);
        }
        if (c != kEvNone) packed += 1ull << (16u * ev_field(c));
    }
    unsigned long long tot;
    block_exclusive_scan<unsigned long long>(packed, &tot);
    if (threadIdx.x < 6u) {
        const uint32_t c = threadIdx.x;
        const bool counted = mine ? (c != 2u && c != 5u) : (c == 2u || c == 5u);
        d.ev_boff[c * nblk + blockIdx.x] = counted ? uint32_t((tot >> (16u * ev_field(c))) & 0xFFFFull) : 0u;
    }
}

// One block: exclusive scan of the (category, block) counts in category-major order, after the
// events already queued; the queue count is clamped to the capacity, the excess counted as dropped.
__global__ void __launch_bounds__(kBlock) k_ev_offsets(DevState d, uint32_t nblk) {
    const uint32_t n = 6u * nblk;
    uint32_t carry = d.cnt->ev_n;
    for (uint32_t start = 0; start < n; start += kEvBlock) {
        const uint32_t i0 = start + threadIdx.x * 4u;
        uint32_t v[4], sum = 0;
        for (int k = 0; k < 4; ++k) {
            v[k] = i0 + k < n ? d.ev_boff[i0 + k] : 0u;
            sum += v[k];
        }
        uint32_t tot;
        uint32_t off = carry + block_exclusive_scan<uint32_t>(sum, &tot);
        for (int k = 0; k < 4; ++k) {
            if (i0 + k < n) d.ev_boff[i0 + k] = off;
            off += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        const uint32_t cap = d.cap_events;
        d.cnt->ev_n = carry < cap ? carry : cap;
        if (carry > cap) d.cnt->ev_dropped += carry - cap;
    }
}

__global__ void __launch_bounds__(kBlock) k_ev_scatter(DevState d, uint32_t nbc, uint32_t nblk, int cur) {
    const bool mine = blockIdx.x < nbc;
    const uint32_t base = (mine ? blockIdx.x : blockIdx.x - nbc) * kEvBlock + threadIdx.x * 4u;
    const uint32_t n = mine ? d.cnt->ev_snap_n[cur] : d.cnt->ev_snap_n[cur ^ 1];
    const uint8_t* codes = mine ? d.ev_code : d.ev_code + d.cap_pairs;
    const unsigned long long* list = mine ? d.ev_snap[cur] : d.ev_snap[cur ^ 1];
    uint32_t c4[4];
    unsigned long long packed = 0;
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t i = base + k;
        c4[k] = i < n ? uint32_t(codes[i]) : kEvNone;
        if (c4[k] != kEvNone) packed += 1ull << (16u * ev_field(c4[k]));
    }
    unsigned long long pre = block_exclusive_scan<unsigned long long>(packed, nullptr);
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t c = c4[k];
        if (c == kEvNone) continue;
        const uint32_t f = ev_field(c);
        const uint32_t pos = d.ev_boff[c * nblk + blockIdx.x] + uint32_t((pre >> (16u * f)) & 0xFFFFull);
        pre += 1ull << (16u * f);
        if (pos >= d.cap_events) continue;
        const unsigned long long key = list[base + k];
        d.ev_out[pos] = make_uint4(d.s_info[key_a(key)].x, d.s_info[key_b(key)].x, c % 3u, c / 3u);
    }
}

}  // namespace

size_t events_sort_bytes(uint32_t cap_pairs) {
    size_t bytes = 0;
    unsigned long long* none = nullptr;
    if (rocprim::radix_sort_keys(nullptr, bytes, none, none, cap_pairs, 0u, kKeyBits) != hipSuccess) return 0;
    return bytes;
}

hipError_t launch_events(const DevState& d, uint32_t n_pairs, uint32_t prev_bound, int cur, hipStream_t st) {
    if (n_pairs) {
        hipLaunchKernelGGL(k_ev_keys, dim3((n_pairs + kBlock - 1) / kBlock), dim3(kBlock), 0, st, d, n_pairs, cur);
        size_t bytes = d.ev_sort_bytes;
        hipError_t e = rocprim::radix_sort_keys(d.ev_sort_tmp, bytes, d.ev_keys, d.ev_snap[cur], n_pairs, 0u, kKeyBits, st);
        if (e != hipSuccess) return e;
    } else {
        hipLaunchKernelGGL(k_ev_empty, dim3(1), dim3(1), 0, st, d, cur);
    }
    if (!n_pairs && !prev_bound) return hipSuccess;
    const uint32_t nbc = (n_pairs + kEvBlock - 1) / kEvBlock, nbp = (prev_bound + kEvBlock - 1) / kEvBlock;
    const uint32_t nblk = nbc + nbp;
    hipLaunchKernelGGL(k_ev_classify, dim3(nblk), dim3(kBlock), 0, st, d, nbc, nblk, cur);
    hipLaunchKernelGGL(k_ev_offsets, dim3(1), dim3(kBlock), 0, st, d, nblk);
    hipLaunchKernelGGL(k_ev_scatter, dim3(nblk), dim3(kBlock), 0, st, d, nbc, nblk, cur);
    return hipSuccess;
}

}  // namespace pxg
