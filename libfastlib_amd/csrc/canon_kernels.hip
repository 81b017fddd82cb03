// Canonical-result WAH kernels (gfx950).
//
// The c0 / d1 / d2 / c1 routines of the reference produce a result whose
// words after compress() are the canonical WAH of the result bits
// (src/bitvector.cpp:2613-2830 and siblings write literal arrays).  By the
// reference's own dispatch these routines run when the operands are dense
// ((|A|w + |B|w) * 31 >= nbits, or an operand is literal-only), so the kernel
// works group by group:
//   1. each block owns CT consecutive 31-bit groups; a compressed operand's
//      covering code words are staged in LDS (found through the chunk
//      index) and expanded, a literal operand is loaded directly;
//   2. the op is applied word-wise;
//   3. canonical compression: a group heads an output word when it is a
//      mixed literal or starts a run of equal all-0/all-1 groups; the word
//      of a run is written by the run's last group.  Heads are counted with
//      a (count, last head) scan -- wave DPP sums, shuffle max, one LDS
//      round per block -- and a wave-parallel look-back between blocks.
#include "dev_common.h"

namespace fbk {
using namespace wah;

constexpr int CT_TPB = 256;
constexpr int CT_ITEMS = 16;
constexpr int CT = CT_TPB * CT_ITEMS;        // groups per block
constexpr uint32_t CHUNKW = 32;

struct Operand {
    const word_t* w;
    const uint32_t* ck;     // chunk index (ignored for literal operands)
    uint32_t n;             // code words
    uint32_t literal;       // one word per group
};

struct CanonArgs {
    Operand a, b;           // b.w == nullptr: single operand
    uint32_t groups;
    int op;
    word_t* out;
    uint32_t* outck;
    uint64_t* status;
    uint32_t* ctl;          // [0] ticket [1] words [2] ones [3] groups
};

__global__ void k_zero_u32_c(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

__global__ void k_zero_u64_c(uint64_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

__device__ __forceinline__ uint32_t cls3(word_t v) { return v == 0 ? 1u : (v == ALLONES ? 2u : 3u); }

// Value of group g of an operand (global lookups; used for tile edges).
__device__ word_t value_at(const Operand& o, uint32_t g) {
    if (o.literal) return o.w[g];
    const uint32_t nc = (o.n + CHUNKW - 1) / CHUNKW;
    const uint32_t c = last_le(o.ck, nc, g);
    uint32_t i = c * CHUNKW, pos = o.ck[c];
    for (;;) {
        const word_t x = o.w[i];
        const uint32_t l = word_len(x);
        if (pos + l > g) return group_value(x);
        pos += l; ++i;
    }
}

// Expands operand `o` over groups [g0, g0+cnt) into vals (combining with the
// existing values by `op` when `combine` is set).
__device__ void expand(const Operand& o, uint32_t g0, uint32_t cnt, word_t* vals, word_t* sw,
                       uint32_t* sp, uint32_t* shs, uint32_t* info, bool combine, int op) {
    const int tid = threadIdx.x;
    if (o.literal) {
        for (uint32_t k = tid; k < cnt; k += CT_TPB) {
            const word_t x = o.w[g0 + k];
            vals[k] = combine ? apply_op(vals[k], x, op) : x;
        }
        __syncthreads();
        return;
    }
    if (tid == 0) {
        const uint32_t nc = (o.n + CHUNKW - 1) / CHUNKW;
        const uint32_t c = last_le(o.ck, nc, g0);
        uint32_t i = c * CHUNKW, pos = o.ck[c];
        for (;;) {
            const uint32_t l = word_len(o.w[i]);
            if (pos + l > g0) break;
            pos += l; ++i;
        }
        info[0] = i; info[1] = pos;
    }
    __syncthreads();
    const uint32_t i0 = info[0], p0 = info[1];
    const uint32_t nw = min((uint32_t)CT + 1u, o.n - i0);
    for (uint32_t k = tid; k < nw; k += CT_TPB) sw[k] = o.w[i0 + k];
    __syncthreads();
    constexpr int IT = (CT + 1 + CT_TPB - 1) / CT_TPB;
    uint32_t len[IT];
    uint32_t s = 0;
    #pragma unroll
    for (int k = 0; k < IT; ++k) {
        const uint32_t j = tid * IT + k;
        len[k] = j < nw ? word_len(sw[j]) : 0u;
        s += len[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_sum<CT_TPB>(s, shs, &tot) + p0;
    #pragma unroll
    for (int k = 0; k < IT; ++k) {
        const uint32_t j = tid * IT + k;
        if (j <= nw) sp[j] = ex;
        ex += len[k];
    }
    __syncthreads();
    // each thread expands its CT_ITEMS consecutive groups
    const uint32_t k0 = tid * CT_ITEMS;
    if (k0 < cnt) {
        uint32_t j = last_le(sp, nw + 1, g0 + k0);
        #pragma unroll
        for (int q = 0; q < CT_ITEMS; ++q) {
            const uint32_t k = k0 + q;
            if (k >= cnt) break;
            while (sp[j + 1] <= g0 + k) ++j;
            const word_t x = group_value(sw[j]);
            vals[k] = combine ? apply_op(vals[k], x, op) : x;
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(CT_TPB) void k_canon_tiles(CanonArgs p) {
    __shared__ word_t s_vals[CT + 2];       // [0] = group before the tile, [CT+1] after
    __shared__ word_t s_sw[CT + 1];         // staged code words, then output staging
    __shared__ uint32_t s_sp[CT + 2];
    __shared__ uint32_t s_shs[CT_TPB / WAVE + 1];
    __shared__ uint32_t s_info[8];
    __shared__ uint32_t s_wc[CT_TPB / WAVE], s_wh[CT_TPB / WAVE];
    __shared__ uint32_t s_tile[4];

    const int tid = threadIdx.x;
    if (tid == 0) s_tile[0] = atomicAdd(&p.ctl[0], 1u);
    __syncthreads();
    const uint32_t tile = s_tile[0];
    const uint32_t g0 = tile * (uint32_t)CT;
    const uint32_t cnt = g0 < p.groups ? min((uint32_t)CT, p.groups - g0) : 0u;
    word_t* vals = s_vals + 1;

    if (cnt) {
        expand(p.a, g0, cnt, vals, s_sw, s_sp, s_shs, s_info, false, p.op);
        if (p.b.w) expand(p.b, g0, cnt, vals, s_sw, s_sp, s_shs, s_info, true, p.op);
    }
    if (tid == 0) {
        // neighbours of the tile (class 0 = none)
        word_t vp = 0xFFFFFFFFu, vn = 0xFFFFFFFFu;
        if (cnt && g0 > 0) {
            vp = value_at(p.a, g0 - 1);
            if (p.b.w) vp = apply_op(vp, value_at(p.b, g0 - 1), p.op);
        }
        if (cnt && g0 + cnt < p.groups) {
            vn = value_at(p.a, g0 + cnt);
            if (p.b.w) vn = apply_op(vn, value_at(p.b, g0 + cnt), p.op);
        }
        s_vals[0] = vp;
        vals[cnt] = vn;          // slot right after the last group of the tile
    }
    __syncthreads();

    // count pass over this thread's groups
    const uint32_t k0 = tid * CT_ITEMS;
    uint32_t c = 0, lh1 = 0;
    #pragma unroll
    for (int q = 0; q < CT_ITEMS; ++q) {
        const uint32_t k = k0 + q;
        if (k < cnt) {
            const word_t v = vals[k];
            const uint32_t cl = cls3(v);
            const word_t pv = vals[(int)k - 1];
            const bool head = (cl == 3u) || (g0 + k == 0) || (pv == 0xFFFFFFFFu) || (cls3(pv) != cl);
            if (head) { ++c; lh1 = g0 + k + 1; }
        }
    }
    // block scan: counts (sum) and last heads (max)
    const int lane = (int)lane_id();
    const int w = tid / WAVE;
    const uint32_t ci = wave_incl_sum(c);
    uint32_t hi = lh1;
    #pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)hi, d, 64);
        if (lane >= d) hi = hi > o ? hi : o;
    }
    if (lane == WAVE - 1) { s_wc[w] = ci; s_wh[w] = hi; }
    __syncthreads();
    uint32_t wc = 0, wh = 0, tc = 0, th = 0;
    for (int i = 0; i < CT_TPB / WAVE; ++i) {
        if (i < w) { wc += s_wc[i]; wh = wh > s_wh[i] ? wh : s_wh[i]; }
        tc += s_wc[i]; th = th > s_wh[i] ? th : s_wh[i];
    }
    uint32_t hx = (uint32_t)__shfl_up((int)hi, 1, 64);
    if (lane == 0) hx = 0;
    const uint32_t t_cnt_ex = wc + ci - c;
    const uint32_t t_lh_ex = wh > hx ? wh : hx;
    if (w == 0) {
        uint32_t ec, eh;
        lookback_ch_wave(p.status, tile, tc, th, ec, eh);
        if (lane == 0) { s_tile[1] = ec; s_tile[2] = eh; }
    }
    __syncthreads();
    const uint32_t base_cnt = s_tile[1];
    const uint32_t base_lh = s_tile[2];
    uint32_t run_cnt = base_cnt + t_cnt_ex;
    uint32_t run_lh1 = base_lh > t_lh_ex ? base_lh : t_lh_ex;

    // write pass: words with slot >= base_cnt go to LDS staging, a run that
    // began in an earlier tile is written straight to global memory
    word_t* stage = s_sw;
    uint32_t ones = 0;
    #pragma unroll
    for (int q = 0; q < CT_ITEMS; ++q) {
        const uint32_t k = k0 + q;
        if (k < cnt) {
            const uint32_t g = g0 + k;
            const word_t v = vals[k];
            const uint32_t cl = cls3(v);
            const word_t pv = vals[(int)k - 1];
            const bool head = (cl == 3u) || (g == 0) || (pv == 0xFFFFFFFFu) || (cls3(pv) != cl);
            if (head) { ++run_cnt; run_lh1 = g + 1; }
            bool emit = false;
            word_t word = v;
            uint32_t headpos = g;
            if (cl == 3u) {
                emit = true;
            } else {
                const word_t nv = vals[k + 1];
                const bool end = (g + 1 == p.groups) || (nv == 0xFFFFFFFFu) || (cls3(nv) != cl);
                if (end) {
                    emit = true;
                    headpos = run_lh1 - 1;
                    const uint32_t len = g - headpos + 1;
                    if (len >= 2) word = (cl == 2u ? HEADER1 : HEADER0) | len;
                }
            }
            if (emit) {
                const uint32_t slot = run_cnt - 1;
                if (slot >= base_cnt) stage[slot - base_cnt] = word;
                else p.out[slot] = word;
                if ((slot & (CHUNKW - 1)) == 0) p.outck[slot / CHUNKW] = headpos;
                ones += (word >= HEADER1) ? (word & MAXCNT) * 31u
                                          : (word < HEADER0 ? (uint32_t)__popc(word) : 0u);
            }
            if (g + 1 == p.groups) {
                p.ctl[1] = run_cnt;
                p.ctl[3] = p.groups;
                p.outck[(run_cnt + CHUNKW - 1) / CHUNKW] = p.groups;
            }
        }
    }
    __syncthreads();
    // coalesced store of the staged words; the tile's trailing open run (it
    // continues into the next tile) is left to the tile that ends it
    if (cnt) {
        const word_t lastv = vals[cnt - 1];
        const word_t nv = vals[cnt];
        const bool open = (cls3(lastv) != 3u) && (nv != 0xFFFFFFFFu) && (cls3(nv) == cls3(lastv));
        // tc == 0: the whole tile lies inside a run headed in an earlier
        // tile, so it owns no slot at all
        const uint32_t nstage = (open && tc > 0u) ? tc - 1u : tc;
        for (uint32_t k = tid; k < nstage; k += CT_TPB) p.out[base_cnt + k] = stage[k];
    }
    ones = wave_sum(ones);
    if (lane == 0 && ones) atomicAdd(&p.ctl[2], ones);
}

hipError_t launch_canon(const word_t* a, const uint32_t* cka, uint32_t na, int alit,
                        const word_t* b, const uint32_t* ckb, uint32_t nb, int blit,
                        uint32_t groups, int op, word_t* out, uint32_t* outck, uint64_t* status,
                        uint32_t* ctl, hipStream_t s) {
    CanonArgs p;
    p.a = Operand{a, cka, na, (uint32_t)(alit ? 1 : 0)};
    p.b = Operand{b, ckb, nb, (uint32_t)(blit ? 1 : 0)};
    p.groups = groups; p.op = op; p.out = out; p.outck = outck; p.status = status; p.ctl = ctl;
    const uint32_t nt = groups ? (groups + CT - 1) / CT : 1u;
    hipLaunchKernelGGL(k_zero_u64_c, dim3(min((nt + 255) / 256, 4096u)), dim3(256), 0, s, status, nt);
    hipLaunchKernelGGL(k_zero_u32_c, dim3(1), dim3(64), 0, s, ctl, 8u);
    hipLaunchKernelGGL(k_canon_tiles, dim3(nt), dim3(CT_TPB), 0, s, p);
    return hipGetLastError();
}

uint32_t canon_tiles(uint32_t groups) { return groups ? (groups + CT - 1) / CT : 1u; }

} // namespace fbk
