// Device-side helpers shared by the gfx950 kernels: wave/block scans built
// on DPP row shifts and broadcasts (64-lane wavefronts), the wave-parallel
// look-back between tiles, and small search routines.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "wah_core.h"

namespace fbk {

constexpr int WAVE = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// An empty use of a loaded value: placed after a batch of loads, it makes
// the compiler issue the whole batch before waiting for any of it.
template <class T>
__device__ __forceinline__ void use_loaded(T x) {
    if constexpr (sizeof(T) == 8) asm volatile("" ::"v"(__builtin_bit_cast(uint64_t, x)));
    else if constexpr (sizeof(T) == 4) asm volatile("" ::"v"(__builtin_bit_cast(uint32_t, x)));
    else asm volatile("" ::"v"((uint32_t)x));
}

// Inclusive prefix sum across a 64-lane wavefront: row_shr 1,2,4,8 inside
// each 16-lane row, then row_bcast 15 / 31 to carry across rows.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false); // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false); // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false); // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false); // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false); // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false); // row_bcast:31
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    v = wave_incl_sum(v);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    // 64-bit sum through two 32-bit passes with carry handled by splitting
    // the low half into 16-bit pieces (no overflow for 64 lanes).
    uint32_t l0 = wave_sum(lo & 0xFFFFu);
    uint32_t l1 = wave_sum(lo >> 16);
    uint32_t h = wave_sum(hi);
    return (uint64_t)l0 + ((uint64_t)l1 << 16) + ((uint64_t)h << 32);
}

// Exclusive block-wide sum.  `sh` needs NWAVES+1 words of LDS.  Returns the
// exclusive prefix of the calling thread; *total receives the block sum.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t* sh, uint32_t* total) {
    constexpr int NW = NT / WAVE;
    const uint32_t w = threadIdx.x / WAVE;
    const uint32_t incl = wave_incl_sum(v);
    if (lane_id() == WAVE - 1) sh[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < NW; ++i) { uint32_t t = sh[i]; sh[i] = acc; acc += t; }
        sh[NW] = acc;
    }
    __syncthreads();
    const uint32_t r = sh[w] + incl - v;
    *total = sh[NW];
    __syncthreads();
    return r;
}

// Look-back between tiles.  Tile t is block t: the dispatcher starts
// blocks in index order, so every predecessor of a tile is resident or
// done (a shared atomic ticket would do the same at ~20 ns per block, the
// cost of serialised same-address atomics).  A
// tile publishes its aggregate (status 1), reads its predecessors 64 at a
// time until an inclusive prefix (status 2) closes the chain, then
// publishes its own inclusive prefix.  Every value a tile needs from
// another travels inside the 64-bit status word, so the loads and stores
// are relaxed agent-scope atomics: release / acquire would write back and
// invalidate the XCD's whole L2 on every step.

// Largest index i in [0, n) with a[i] <= key (a sorted, a[0] <= key).
template <class T>
__device__ __forceinline__ uint32_t last_le(const T* a, uint32_t n, uint32_t key) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t o = (uint32_t)__shfl_xor((int)v, d, 64);
        v = v > o ? v : o;
    }
    return v;
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

// ---------------------------------------------------------------------------
// Run summaries across a wavefront (shuffles; lane order = position order).
__device__ __forceinline__ wah::Summary shfl_summary_up(const wah::Summary& x, int d) {
    wah::Summary r;
    r.count = __shfl_up(x.count, d, 64);
    r.tlen = __shfl_up(x.tlen, d, 64);
    r.first = __shfl_up(x.first, d, 64);
    r.trail = __shfl_up(x.trail, d, 64);
    r.uniform = __shfl_up(x.uniform, d, 64);
    r.trep = __shfl_up(x.trep, d, 64);
    return r;
}

__device__ __forceinline__ wah::Summary shfl_summary_down(const wah::Summary& x, int d) {
    wah::Summary r;
    r.count = __shfl_down(x.count, d, 64);
    r.tlen = __shfl_down(x.tlen, d, 64);
    r.first = __shfl_down(x.first, d, 64);
    r.trail = __shfl_down(x.trail, d, 64);
    r.uniform = __shfl_down(x.uniform, d, 64);
    r.trep = __shfl_down(x.trep, d, 64);
    return r;
}

__device__ __forceinline__ wah::Summary shfl_summary_idx(const wah::Summary& x, int src) {
    wah::Summary r;
    r.count = __shfl(x.count, src, 64);
    r.tlen = __shfl(x.tlen, src, 64);
    r.first = __shfl(x.first, src, 64);
    r.trail = __shfl(x.trail, src, 64);
    r.uniform = __shfl(x.uniform, src, 64);
    r.trep = __shfl(x.trep, src, 64);
    return r;
}

// Inclusive scan of summaries over the lanes of a wavefront.
__device__ __forceinline__ wah::Summary wave_incl_summary(wah::Summary x) {
    const int lane = (int)lane_id();
    #pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const wah::Summary o = shfl_summary_up(x, d);
        if (lane >= d) x = wah::combine(o, x);
    }
    return x;
}

// Exclusive block scan of summaries: wave scans + one LDS round.
// `wsum` needs NT/64 Summary slots.
template <int NT>
__device__ __forceinline__ wah::Summary block_excl_summary2(const wah::Summary& v, wah::Summary* wsum,
                                                            wah::Summary* agg) {
    constexpr int NW = NT / WAVE;
    const int lane = (int)lane_id();
    const int w = threadIdx.x / WAVE;
    const wah::Summary incl = wave_incl_summary(v);
    wah::Summary excl = shfl_summary_up(incl, 1);
    if (lane == 0) excl = wah::summary_empty();
    if (lane == WAVE - 1) wsum[w] = incl;
    __syncthreads();
    wah::Summary pre = wah::summary_empty();
    for (int i = 0; i < w; ++i) pre = wah::combine(pre, wsum[i]);
    wah::Summary all = pre;
    for (int i = w; i < NW; ++i) all = wah::combine(all, wsum[i]);
    *agg = all;
    __syncthreads();
    return wah::combine(pre, excl);
}

// Wave-parallel look-back over run summaries.  Called by all 64 lanes of
// one wavefront; lane k reads the status of tile (base - k).  Returns the
// exclusive prefix (valid in every lane).
__device__ __forceinline__ wah::Summary lookback_summary_wave(uint64_t* status, uint32_t tile,
                                                              const wah::Summary& agg) {
    const int lane = (int)lane_id();
    if (tile == 0) {
        if (lane == 0)
            __hip_atomic_store(&status[0], wah::pack(agg, 2), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return wah::summary_empty();
    }
    if (lane == 0)
        __hip_atomic_store(&status[tile], wah::pack(agg, 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    wah::Summary acc = wah::summary_empty();
    int64_t base = (int64_t)tile - 1;
    for (;;) {
        const int64_t j = base - lane;
        uint64_t s = (j >= 0) ? __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : wah::pack(wah::summary_empty(), 2);
        const uint32_t st = wah::status_of(s);
        const uint64_t zero = ballot64(st == 0);
        const uint64_t pref = ballot64(st == 2);
        const uint64_t need = pref ? ((pref & (~pref + 1)) << 1) - 1 : ~0ull;
        if (zero & need) { __builtin_amdgcn_s_sleep(1); continue; }
        wah::Summary x = ((need >> lane) & 1ull) ? wah::unpack(s) : wah::summary_empty();
        // combine far (high lane) to near (lane 0)
        #pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const wah::Summary o = shfl_summary_down(x, d);
            if (lane + d < 64) x = wah::combine(o, x);
        }
        const wah::Summary win = shfl_summary_idx(x, 0);
        acc = wah::combine(win, acc);
        if (pref) break;
        base -= 64;
    }
    if (lane == 0)
        __hip_atomic_store(&status[tile], wah::pack(wah::combine(acc, agg), 2), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    return acc;
}

// (count, last head + 1) pairs for canonical compression: count adds,
// last-head takes the maximum.  Packed [63:62] status, [61:31] count,
// [30:0] last head + 1 (0 = none).
__device__ __forceinline__ uint64_t pack_ch(uint32_t cnt, uint32_t lh1, uint64_t st) {
    return (st << 62) | ((uint64_t)(cnt & 0x7FFFFFFFu) << 31) | (uint64_t)(lh1 & 0x7FFFFFFFu);
}

// Wave-parallel look-back for (count, last head): all 64 lanes of a wave.
__device__ __forceinline__ void lookback_ch_wave(uint64_t* status, uint32_t tile, uint32_t acnt,
                                                 uint32_t alh1, uint32_t& ecnt, uint32_t& elh1) {
    const int lane = (int)lane_id();
    if (tile == 0) {
        if (lane == 0)
            __hip_atomic_store(&status[0], pack_ch(acnt, alh1, 2), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ecnt = 0; elh1 = 0;
        return;
    }
    if (lane == 0)
        __hip_atomic_store(&status[tile], pack_ch(acnt, alh1, 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t c = 0, h = 0;
    int64_t base = (int64_t)tile - 1;
    for (;;) {
        const int64_t j = base - lane;
        uint64_t s = (j >= 0) ? __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : pack_ch(0, 0, 2);
        const uint32_t st = (uint32_t)(s >> 62);
        const uint64_t zero = ballot64(st == 0);
        const uint64_t pref = ballot64(st == 2);
        const uint64_t need = pref ? ((pref & (~pref + 1)) << 1) - 1 : ~0ull;
        if (zero & need) { __builtin_amdgcn_s_sleep(1); continue; }
        const bool use = (need >> lane) & 1ull;
        c += wave_sum(use ? (uint32_t)((s >> 31) & 0x7FFFFFFFu) : 0u);
        const uint32_t hm = wave_max(use ? (uint32_t)(s & 0x7FFFFFFFu) : 0u);
        h = h > hm ? h : hm;
        if (pref) break;
        base -= 64;
    }
    ecnt = c; elh1 = h;
    if (lane == 0)
        __hip_atomic_store(&status[tile], pack_ch(c + acnt, h > alh1 ? h : alh1, 2), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Run summary of one wave's lanes in lane order, valid in every lane.
__device__ __forceinline__ wah::Summary wave_total_summary(const wah::Summary& x) {
    return shfl_summary_idx(wave_incl_summary(x), 63);
}

// Spins until the summary entry p carries `tag`, then returns it.
__device__ __forceinline__ wah::Summary wait_summary(const uint64_t* p, uint32_t tag) {
    uint64_t v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (wah::status_of(v) != tag) {
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return wah::unpack(v);
}

// Two-level exclusive prefix of run summaries for kernels whose tile t is
// block t (blocks start in index order).  Tile t publishes its aggregate in
// aggs[t]; the last tile of every segment of 64 tiles publishes the
// segment's aggregate in segs[t / 64] as soon as the tiles before it in the
// segment have published theirs, before it waits for anything else.  A
// tile then combines the aggregates of the tiles before it in its segment
// and of the segments before its own: two rounds of independent loads,
// where a walk back window by window takes a round trip per 64 tiles while
// every tile is still waiting.  Entries carry `tag` (1 or 2): callers that
// alternate the tag between launches on the same buffers, all of one tile
// count, start from zeroed buffers and never clear them again.  Called by
// all 64 lanes of one wave; the result is valid in every lane.
__device__ __forceinline__ wah::Summary prefix_summary_seg(uint64_t* aggs, uint64_t* segs, uint32_t tile,
                                                           uint32_t tag, const wah::Summary& agg) {
    const uint32_t lane = lane_id();
    const uint32_t s = tile >> 6, r = tile & 63u;
    if (lane == 0) __hip_atomic_store(&aggs[tile], wah::pack(agg, tag), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const wah::Summary mine = lane < r ? wait_summary(&aggs[(s << 6) + lane], tag) : wah::summary_empty();
    const wah::Summary in_seg = wave_total_summary(mine);
    if (r == 63u && lane == 0)
        __hip_atomic_store(&segs[s], wah::pack(wah::combine(in_seg, agg), tag), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    wah::Summary before = wah::summary_empty();
    for (uint32_t k0 = 0; k0 < s; k0 += 64u) {
        const uint32_t j = k0 + lane;
        const wah::Summary x = j < s ? wait_summary(&segs[j], tag) : wah::summary_empty();
        before = wah::combine(before, wave_total_summary(x));
    }
    return wah::combine(before, in_seg);
}

} // namespace fbk
