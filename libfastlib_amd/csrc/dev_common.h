// Device-side helpers shared by the gfx950 kernels: wave/block scans built
// on DPP row shifts and broadcasts (64-lane wavefronts), the single-pass
// look-back between tiles, and small search routines.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "wah_core.h"

namespace fbk {

constexpr int WAVE = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

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

// Block-wide exclusive scan of run summaries (LDS, Hillis-Steele).  `buf`
// needs 2*NT Summary slots.  Returns the exclusive prefix of the calling
// thread and writes the block aggregate to *agg.
template <int NT>
__device__ __forceinline__ wah::Summary block_excl_summary(const wah::Summary& v, wah::Summary* buf,
                                                           wah::Summary* agg) {
    const int t = threadIdx.x;
    wah::Summary* in = buf;
    wah::Summary* out = buf + NT;
    in[t] = v;
    __syncthreads();
    for (int d = 1; d < NT; d <<= 1) {
        wah::Summary x = in[t];
        if (t >= d) x = wah::combine(in[t - d], x);
        out[t] = x;
        __syncthreads();
        wah::Summary* tmp = in; in = out; out = tmp;
    }
    const wah::Summary excl = (t == 0) ? wah::summary_empty() : in[t - 1];
    *agg = in[NT - 1];
    __syncthreads();
    return excl;
}

// Single-pass look-back between tiles (tile ids handed out in launch order
// by an atomic ticket, so every predecessor is resident).  Called by one
// thread.  Publishes this tile's aggregate, walks back over predecessors
// until an inclusive prefix is found, publishes the inclusive prefix and
// returns the exclusive one.
__device__ __forceinline__ wah::Summary lookback(uint64_t* status, uint32_t tile,
                                                 const wah::Summary& agg) {
    if (tile == 0) {
        __hip_atomic_store(&status[0], wah::pack(agg, 2), __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
        return wah::summary_empty();
    }
    __hip_atomic_store(&status[tile], wah::pack(agg, 1), __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_AGENT);
    wah::Summary acc = wah::summary_empty();
    int64_t j = (int64_t)tile - 1;
    while (j >= 0) {
        uint64_t s;
        do {
            s = __hip_atomic_load(&status[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            if (wah::status_of(s) == 0) __builtin_amdgcn_s_sleep(1);
        } while (wah::status_of(s) == 0);
        acc = wah::combine(wah::unpack(s), acc);
        if (wah::status_of(s) == 2) break;
        --j;
    }
    __hip_atomic_store(&status[tile], wah::pack(wah::combine(acc, agg), 2), __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_AGENT);
    return acc;
}

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

} // namespace fbk
