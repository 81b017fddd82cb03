// gfx950 kernels of the bitvector operations beside the operator kernels:
//
//  k_count_and      count(mask): ones of a AND b per tile of groups, the
//                   result itself never written
//  k_lit_op         a OP b over two literal arrays (the literal results of
//                   the out-of-place c0 and d2 routines)
//  k_c1_heads       value and word-start flag of every result group of the
//                   *_c1 loops (one operand compressed, one literal)
//  k_head_pos,      result words from the word-start flags and their
//  k_head_words     exclusive scan
//  k_shift_groups   the literal groups of a bit sequence cut out of a
//                   bitvector at another offset (erase)
//
// A compressed operand is read through its chunk index: a block stages the
// code words covering its tile of groups and their start groups in LDS,
// and each thread finds a group's word by binary search.
#include "dev_common.h"
#include "launch.h"

namespace fbk {
using namespace wah;

constexpr int AX_TPB = 256;
constexpr int AX_TILE = 2048;          // groups per block
constexpr uint32_t AX_CHUNK = 32;      // code words per chunk-index entry

// Code words covering groups [g0, g0 + AX_TILE) into sw[0, cnt) and their
// start groups into sp[0, cnt] (sp[cnt] = end of the last one); info[0] /
// info[1] receive the first staged word's index and start.  Called by every
// thread of the block.
__device__ uint32_t stage_words(const word_t* __restrict__ w, const uint32_t* __restrict__ ck, uint32_t n,
                                uint32_t g0, word_t* sw, uint32_t* sp, uint32_t* shs, uint32_t* info) {
    if (threadIdx.x == 0) {
        const uint32_t nchunks = (n + AX_CHUNK - 1) / AX_CHUNK;
        const uint32_t c = last_le(ck, nchunks, g0);
        uint32_t i = c * AX_CHUNK, pos = ck[c];
        while (i < n) {
            const uint32_t l = word_len(w[i]);
            if (pos + l > g0) break;
            pos += l; ++i;
        }
        info[0] = i; info[1] = pos;
    }
    __syncthreads();
    const uint32_t i0 = info[0], p0 = info[1];
    const uint32_t cnt = min((uint32_t)AX_TILE + 1u, n - i0);
    for (uint32_t k = threadIdx.x; k < cnt; k += AX_TPB) sw[k] = w[i0 + k];
    __syncthreads();
    constexpr int IT = (AX_TILE + 1 + AX_TPB - 1) / AX_TPB;
    uint32_t l[IT];
    uint32_t s = 0;
    #pragma unroll
    for (int k = 0; k < IT; ++k) {
        const uint32_t j = threadIdx.x * IT + k;
        l[k] = j < cnt ? word_len(sw[j]) : 0u;
        s += l[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_sum<AX_TPB>(s, shs, &tot) + p0;
    #pragma unroll
    for (int k = 0; k < IT; ++k) {
        const uint32_t j = threadIdx.x * IT + k;
        if (j <= cnt) sp[j] = ex;
        ex += l[k];
    }
    __syncthreads();
    return cnt;
}

__device__ __forceinline__ uint32_t block_total(uint32_t v, uint32_t* shs) {
    v = wave_sum(v);
    if (lane_id() == 0) shs[threadIdx.x / WAVE] = v;
    __syncthreads();
    uint32_t t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < AX_TPB / WAVE; ++i) t += shs[i];
    return t;
}

// ---------------------------------------------------------------------------
// count(mask): part[block] = ones of a AND b over the block's groups; an
// operand with lit != 0 is one literal word per group.
__global__ __launch_bounds__(AX_TPB) void k_count_and(const word_t* __restrict__ a, const uint32_t* __restrict__ cka,
                                                      uint32_t na, int alit, const word_t* __restrict__ b,
                                                      const uint32_t* __restrict__ ckb, uint32_t nb, int blit,
                                                      uint32_t groups, uint32_t* __restrict__ part) {
    __shared__ word_t swa[AX_TILE + 1], swb[AX_TILE + 1];
    __shared__ uint32_t spa[AX_TILE + 2], spb[AX_TILE + 2];
    __shared__ uint32_t shs[AX_TPB / WAVE + 1];
    __shared__ uint32_t info[4];
    const uint32_t g0 = blockIdx.x * AX_TILE;
    const uint32_t g1 = min(g0 + (uint32_t)AX_TILE, groups);
    const uint32_t cnta = alit ? 0u : stage_words(a, cka, na, g0, swa, spa, shs, info);
    const uint32_t cntb = blit ? 0u : stage_words(b, ckb, nb, g0, swb, spb, shs, info + 2);
    uint32_t ones = 0;
    for (uint32_t g = g0 + threadIdx.x; g < g1; g += AX_TPB) {
        const word_t va = alit ? a[g] : group_value(swa[last_le(spa, cnta, g)]);
        const word_t vb = blit ? b[g] : group_value(swb[last_le(spb, cntb, g)]);
        ones += (uint32_t)__popc(va & vb);
    }
    const uint32_t t = block_total(ones, shs);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(AX_TPB) void k_lit_op(const word_t* __restrict__ a, const word_t* __restrict__ b,
                                                   uint32_t n, int op, word_t* __restrict__ out) {
    for (uint32_t i = blockIdx.x * AX_TPB + threadIdx.x; i < n; i += gridDim.x * AX_TPB)
        out[i] = apply_op(a[i], b[i], op);
}

// ---------------------------------------------------------------------------
// The *_c1 loops (src/bitvector.cpp:2545, 2981, 3354, 3748, 3821) walk the
// compressed operand x over the literal operand y: a literal word of x
// appends x OP y (append_active); a fill whose result is constant appends
// one fill token over its span (append_counter, or append_active for one
// group); any other fill pushes the result of each group raw.  A group
// therefore starts a result word unless it continues a fill token, or its
// token merges into the word before it: a token that is not a raw push,
// of the same all-zero / all-one class as the previous group.  A run of two
// or more merged groups is one fill word; a lone group is its own value
// (raw pushes are literals here).
enum : int { RK_AND = 0, RK_OR = 1, RK_XOR = 2, RK_MINUS = 3, RK_MINUSX = 4 };

__device__ __forceinline__ word_t c1_value(word_t w, word_t y, int rk) {
    const word_t xv = group_value(w);
    switch (rk) {
    case RK_AND: return xv & y;
    case RK_OR: return xv | y;
    case RK_XOR: return xv ^ y;
    case RK_MINUS: return xv & ~y & ALLONES;
    default: return y & ~xv & ALLONES;        // minus_c1x: y - x
    }
}

__device__ __forceinline__ bool c1_copies(word_t w, int rk) {
    const bool one = w >= HEADER1;
    switch (rk) {
    case RK_AND: return one;
    case RK_OR: return !one;
    case RK_XOR: return true;
    case RK_MINUS: return one;
    default: return !one;
    }
}

__global__ __launch_bounds__(AX_TPB) void k_c1_heads(const word_t* __restrict__ x, const uint32_t* __restrict__ ckx,
                                                     uint32_t nx, const word_t* __restrict__ y, uint32_t groups,
                                                     int rk, word_t* __restrict__ V, uint32_t* __restrict__ H) {
    __shared__ word_t sw[AX_TILE + 1];
    __shared__ uint32_t sp[AX_TILE + 2];
    __shared__ word_t vt[AX_TILE + 1];          // vt[0]: the group before the tile
    __shared__ uint32_t shs[AX_TPB / WAVE + 1];
    __shared__ uint32_t info[2];
    const uint32_t g0 = blockIdx.x * AX_TILE;
    const uint32_t g1 = min(g0 + (uint32_t)AX_TILE, groups);
    const uint32_t cnt = stage_words(x, ckx, nx, g0, sw, sp, shs, info);
    if (threadIdx.x == 0 && g0 > 0) {
        // the word holding group g0 - 1: the first staged word, or the one
        // before it when that starts at g0
        const word_t w = sp[0] <= g0 - 1 ? sw[0] : x[info[0] - 1];
        vt[0] = c1_value(w, y[g0 - 1], rk);
    }
    for (uint32_t g = g0 + threadIdx.x; g < g1; g += AX_TPB) {
        const word_t w = sw[last_le(sp, cnt, g)];
        const word_t v = c1_value(w, y[g], rk);
        vt[1 + g - g0] = v;
        V[g] = v;
    }
    __syncthreads();
    for (uint32_t g = g0 + threadIdx.x; g < g1; g += AX_TPB) {
        const uint32_t j = last_le(sp, cnt, g);
        const word_t w = sw[j];
        bool start = true, merge = true;
        if (w > ALLONES) {
            if (c1_copies(w, rk)) merge = false;
            else start = g == sp[j];
        }
        const word_t v = vt[1 + g - g0];
        const uint32_t c = class_of_value(v);
        const bool joins = g > 0 && merge && c != C_MIXED && c == class_of_value(vt[g - g0]);
        H[g] = (start && !joins) ? 1u : 0u;
    }
}

// HP[O[g]] = g for every word start g; HP[total] = n.
__global__ __launch_bounds__(AX_TPB) void k_head_pos(const uint32_t* __restrict__ H, const uint32_t* __restrict__ O,
                                                     uint32_t n, uint32_t* __restrict__ HP) {
    for (uint32_t g = blockIdx.x * AX_TPB + threadIdx.x; g < n; g += gridDim.x * AX_TPB) {
        if (H[g]) HP[O[g]] = g;
        if (g + 1 == n) HP[O[g] + H[g]] = n;
    }
}

__global__ __launch_bounds__(AX_TPB) void k_head_words(const word_t* __restrict__ V, const uint32_t* __restrict__ HP,
                                                       uint32_t nw, word_t* __restrict__ out) {
    for (uint32_t o = blockIdx.x * AX_TPB + threadIdx.x; o < nw; o += gridDim.x * AX_TPB) {
        const uint32_t g = HP[o], len = HP[o + 1] - g;
        const word_t v = V[g];
        out[o] = len >= 2 ? ((v == ALLONES ? HEADER1 : HEADER0) | len) : v;
    }
}

// ---------------------------------------------------------------------------
// erase: group q of the sequence S = bits [p0, p0 + ind) ++ bits [j, ...) of
// a bitvector whose literal groups (the active bits as one more group,
// left-aligned, then a zero word) are L; nS bits in all, bits past nS zero.
__device__ __forceinline__ word_t bits31(const word_t* L, uint32_t pos) {
    const uint32_t w = pos / 31u, o = pos % 31u;
    const word_t hi = L[w] << o;
    const word_t lo = o ? (L[w + 1] >> (31u - o)) : 0u;
    return (hi | lo) & ALLONES;
}

__global__ __launch_bounds__(AX_TPB) void k_shift_groups(const word_t* __restrict__ L, uint32_t p0, uint32_t ind,
                                                         uint32_t j, uint32_t nS, word_t* __restrict__ V) {
    const uint32_t ng = (nS + 30u) / 31u;
    for (uint32_t q = blockIdx.x * AX_TPB + threadIdx.x; q < ng; q += gridDim.x * AX_TPB) {
        const uint32_t k0 = 31u * q;
        word_t v;
        if (k0 + 31u <= ind) {
            v = bits31(L, p0 + k0);
        } else if (k0 >= ind) {
            v = bits31(L, j + (k0 - ind));
        } else {
            const uint32_t m = ind - k0;                       // 1..30 bits from the head
            const word_t keep = ALLONES & ~((1u << (31u - m)) - 1u);
            v = (bits31(L, p0 + k0) & keep) | (bits31(L, j) >> m);
        }
        if (k0 + 31u > nS) v &= ALLONES & ~((1u << (k0 + 31u - nS)) - 1u);
        V[q] = v;
    }
}

// out[0] / out[1]: index and start group of the code word holding group G
// (n and the total groups when G is past the end).
__global__ void k_word_at(const word_t* __restrict__ w, const uint32_t* __restrict__ ck, uint32_t n, uint32_t G,
                          uint32_t* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t nchunks = (n + AX_CHUNK - 1) / AX_CHUNK;
    const uint32_t c = n ? last_le(ck, nchunks, G) : 0u;
    uint32_t i = c * AX_CHUNK, pos = n ? ck[c] : 0u;
    while (i < n) {
        const uint32_t l = word_len(w[i]);
        if (pos + l > G) break;
        pos += l; ++i;
    }
    out[0] = i; out[1] = pos;
}

// ---------------------------------------------------------------------------
static inline uint32_t ax_cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
static inline uint32_t ax_grid(uint32_t n) { return min(ax_cdiv(n ? n : 1u, AX_TPB), 8192u); }

uint32_t count_and_blocks(uint32_t groups) { return groups ? ax_cdiv(groups, AX_TILE) : 1u; }

hipError_t launch_count_and(const word_t* a, const uint32_t* cka, uint32_t na, int alit, const word_t* b,
                            const uint32_t* ckb, uint32_t nb, int blit, uint32_t groups, uint32_t* part,
                            hipStream_t s) {
    if (!groups) return hipSuccess;
    hipLaunchKernelGGL(k_count_and, dim3(count_and_blocks(groups)), dim3(AX_TPB), 0, s, a, cka, na, alit, b, ckb,
                       nb, blit, groups, part);
    return hipGetLastError();
}

hipError_t launch_lit_op(const word_t* a, const word_t* b, uint32_t n, int op, word_t* out, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_lit_op, dim3(ax_grid(n)), dim3(AX_TPB), 0, s, a, b, n, op, out);
    return hipGetLastError();
}

hipError_t launch_c1_heads(const word_t* x, const uint32_t* ckx, uint32_t nx, const word_t* y, uint32_t groups,
                           int rk, word_t* V, uint32_t* H, hipStream_t s) {
    if (!groups) return hipSuccess;
    hipLaunchKernelGGL(k_c1_heads, dim3(ax_cdiv(groups, AX_TILE)), dim3(AX_TPB), 0, s, x, ckx, nx, y, groups, rk,
                       V, H);
    return hipGetLastError();
}

hipError_t launch_head_words(const word_t* V, const uint32_t* H, const uint32_t* O, uint32_t n, uint32_t nw,
                             uint32_t* HP, word_t* out, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_head_pos, dim3(ax_grid(n)), dim3(AX_TPB), 0, s, H, O, n, HP);
    hipLaunchKernelGGL(k_head_words, dim3(ax_grid(nw)), dim3(AX_TPB), 0, s, V, (const uint32_t*)HP, nw, out);
    return hipGetLastError();
}

hipError_t launch_word_at(const word_t* w, const uint32_t* ck, uint32_t n, uint32_t G, uint32_t* out,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_word_at, dim3(1), dim3(64), 0, s, w, ck, n, G, out);
    return hipGetLastError();
}

hipError_t launch_shift_groups(const word_t* L, uint32_t p0, uint32_t ind, uint32_t j, uint32_t nS, word_t* V,
                               hipStream_t s) {
    const uint32_t ng = (nS + 30u) / 31u;
    if (!ng) return hipSuccess;
    hipLaunchKernelGGL(k_shift_groups, dim3(ax_grid(ng)), dim3(AX_TPB), 0, s, L, p0, ind, j, nS, V);
    return hipGetLastError();
}

} // namespace fbk
