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
    const uint2* ts;        // tile starts: {word containing group t*CT, its start group}
};

struct CanonArgs {
    Operand a, b;           // b.w == nullptr: single operand
    uint32_t groups;
    int op;
    word_t* out;
    uint32_t* outck;
    uint64_t* status;
    uint32_t* ctl;          // [0] ticket [1] words [3] groups [4],[5] edge words
};

__global__ void k_zero_u32_c(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

__global__ void k_zero_u64_c(uint64_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

// Tile starts of a compressed operand from its chunk index: the thread of
// chunk j (code words [32j, 32j+32), groups [ck[j], ck[j+1])) resolves the
// tiles whose first group falls in its span by walking those 32 words.
// ts[nt] = {n, groups} closes the table.
__global__ __launch_bounds__(256) void k_tile_starts(const word_t* __restrict__ w, const uint32_t* __restrict__ ck,
                                                     uint32_t n, uint32_t nt, uint2* __restrict__ ts) {
    const uint32_t nc = (n + CHUNKW - 1) / CHUNKW;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > nc) return;
    if (j == nc) { ts[nt] = make_uint2(n, ck[nc]); return; }
    const uint32_t c0 = ck[j], c1 = ck[j + 1];
    uint32_t t = (c0 + CT - 1) / CT;                      // first tile start >= c0
    if ((uint64_t)t * CT >= c1 || t >= nt) return;
    uint32_t i = j * CHUNKW, pos = c0;
    const uint32_t e = min(i + CHUNKW, n);
    for (; i < e && t < nt; ++i) {
        const uint32_t l = word_len(w[i]);
        while (t < nt && (uint64_t)t * CT < (uint64_t)pos + l) {
            ts[t] = make_uint2(i, pos);
            ++t;
        }
        pos += l;
    }
}

__device__ __forceinline__ uint32_t cls3(word_t v) { return v == 0 ? 1u : (v == ALLONES ? 2u : 3u); }

// Expands operand `o` over the tile's groups [g0, g0+cnt) into vals,
// combining with the values there by `op` when `combine` is set, and
// returns (to thread 0) the operand's values of groups g0-1 and g0+cnt
// (0xFFFFFFFF where outside the bitvector).
__device__ void expand(const Operand& o, uint32_t tile, uint32_t g0, uint32_t cnt, uint32_t groups,
                       word_t* vals, word_t* sw, uint32_t* sp, uint32_t* shs, bool combine, int op,
                       word_t& vprev, word_t& vnext) {
    const int tid = threadIdx.x;
    const uint32_t g1 = g0 + cnt;
    if (o.literal) {
        // 16 independent coalesced loads per thread, then LDS
        word_t x[CT_ITEMS];
        #pragma unroll
        for (int r = 0; r < CT_ITEMS; ++r) {
            const uint32_t k = (uint32_t)r * CT_TPB + tid;
            x[r] = k < cnt ? o.w[g0 + k] : 0u;
        }
        #pragma unroll
        for (int r = 0; r < CT_ITEMS; ++r) {
            const uint32_t k = (uint32_t)r * CT_TPB + tid;
            if (k < cnt) vals[k] = combine ? apply_op(vals[k], x[r], op) : x[r];
        }
        if (tid == 0) {
            vprev = g0 > 0 ? o.w[g0 - 1] : 0xFFFFFFFFu;
            vnext = g1 < groups ? o.w[g1] : 0xFFFFFFFFu;
        }
        __syncthreads();
        return;
    }
    const uint2 st = o.ts[tile];
    const uint32_t i0 = st.x, p0 = st.y;
    // words covering the tile: i0 .. the word holding group g1 - 1
    const uint32_t ilast = min(o.ts[tile + 1].x, o.n - 1u);
    const uint32_t nw = ilast + 1u - i0;                   // <= CT + 1
    constexpr int IT = (CT + 1 + CT_TPB - 1) / CT_TPB;
    {
        word_t x[IT];
        #pragma unroll
        for (int r = 0; r < IT; ++r) {
            const uint32_t k = (uint32_t)r * CT_TPB + tid;
            x[r] = k < nw ? o.w[i0 + k] : 0u;
        }
        #pragma unroll
        for (int r = 0; r < IT; ++r) {
            const uint32_t k = (uint32_t)r * CT_TPB + tid;
            if (k < nw) sw[k] = x[r];
        }
    }
    if (tid == 0) {
        // group g0 - 1: in word i0 when it started before g0, else the word before
        vprev = g0 == 0 ? 0xFFFFFFFFu : (p0 < g0 ? 0u : group_value(o.w[i0 - 1]));
        vnext = g1 < groups ? group_value(o.w[o.ts[tile + 1].x]) : 0xFFFFFFFFu;
    }
    __syncthreads();
    if (tid == 0 && g0 > 0 && p0 < g0) vprev = group_value(sw[0]);
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
    __shared__ uint32_t s_wc[CT_TPB / WAVE], s_wh[CT_TPB / WAVE];
    __shared__ uint32_t s_tile[4];

    const int tid = threadIdx.x;
    if (tid == 0) s_tile[0] = atomicAdd(&p.ctl[0], 1u);
    __syncthreads();
    const uint32_t tile = s_tile[0];
    const uint32_t g0 = tile * (uint32_t)CT;
    const uint32_t cnt = g0 < p.groups ? min((uint32_t)CT, p.groups - g0) : 0u;
    word_t* vals = s_vals + 1;

    word_t vpa = 0xFFFFFFFFu, vna = 0xFFFFFFFFu, vpb = 0xFFFFFFFFu, vnb = 0xFFFFFFFFu;
    if (cnt) {
        expand(p.a, tile, g0, cnt, p.groups, vals, s_sw, s_sp, s_shs, false, p.op, vpa, vna);
        if (p.b.w) expand(p.b, tile, g0, cnt, p.groups, vals, s_sw, s_sp, s_shs, true, p.op, vpb, vnb);
    }
    if (tid == 0) {
        // neighbours of the tile (0xFFFFFFFF = none)
        word_t vp = vpa, vn = vna;
        if (p.b.w) {
            if (vp != 0xFFFFFFFFu) vp = apply_op(vpa, vpb, p.op);
            if (vn != 0xFFFFFFFFu) vn = apply_op(vna, vnb, p.op);
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
}

hipError_t launch_canon(const word_t* a, const uint32_t* cka, uint32_t na, int alit,
                        const word_t* b, const uint32_t* ckb, uint32_t nb, int blit,
                        uint32_t groups, int op, word_t* out, uint32_t* outck, uint64_t* status,
                        uint32_t* ctl, uint2* ts, hipStream_t s) {
    CanonArgs p;
    const uint32_t nt = groups ? (groups + CT - 1) / CT : 1u;
    p.a = Operand{a, cka, na, (uint32_t)(alit ? 1 : 0), ts};
    p.b = Operand{b, ckb, nb, (uint32_t)(blit ? 1 : 0), ts + (nt + 1)};
    p.groups = groups; p.op = op; p.out = out; p.outck = outck; p.status = status; p.ctl = ctl;
    if (!alit && na > 0) {
        const uint32_t nc = (na + CHUNKW - 1) / CHUNKW;
        hipLaunchKernelGGL(k_tile_starts, dim3(nc / 256 + 1), dim3(256), 0, s, a, cka, na, nt, ts);
    }
    if (b && !blit && nb > 0) {
        const uint32_t nc = (nb + CHUNKW - 1) / CHUNKW;
        hipLaunchKernelGGL(k_tile_starts, dim3(nc / 256 + 1), dim3(256), 0, s, b, ckb, nb, nt, ts + (nt + 1));
    }
    hipLaunchKernelGGL(k_zero_u64_c, dim3(min((nt + 255) / 256, 4096u)), dim3(256), 0, s, status, nt);
    hipLaunchKernelGGL(k_zero_u32_c, dim3(1), dim3(64), 0, s, ctl, 8u);
    hipLaunchKernelGGL(k_canon_tiles, dim3(nt), dim3(CT_TPB), 0, s, p);
    return hipGetLastError();
}

uint32_t canon_tiles(uint32_t groups) { return groups ? (groups + CT - 1) / CT : 1u; }

} // namespace fbk
