// Canonical-result WAH kernels (gfx950).
//
// The c0 / d1 / d2 / c1 routines of the reference produce a result whose
// words after compress() are the canonical WAH of the result bits
// (src/bitvector.cpp:2613-2830 and siblings write literal arrays).  By the
// reference's own dispatch these routines run when the operands are dense
// ((|A|w + |B|w) * 31 >= nbits, or an operand is literal-only), so the kernel
// works group by group.  Block b owns tile b of CT consecutive 31-bit
// groups; thread t of a tile owns groups
// g0 + q*CT_TPB + t (q < CT_ITEMS), so every load and store is coalesced:
//   1. a literal operand is loaded directly; a compressed operand's
//      covering code words are staged in LDS, their start groups scanned,
//      and each group finds its word through a max-scan of word marks;
//   2. the op is applied word-wise;
//   3. canonical compression: a group heads an output word when it is a
//      mixed literal or starts a run of equal all-0/all-1 groups; the word
//      of a run is written by the run's last group.  Heads are counted by
//      wave ballots per (slice q, wave) segment, segment totals meet in LDS,
//      and a wave-parallel look-back carries (count, last head) between
//      tiles.  A tile waits only for tiles of lower block index, which the
//      dispatcher has started before it (in-order dispatch), so there is no
//      shared ticket counter: same-address atomics serialise (~20 ns each).
#include "dev_common.h"
#include <mutex>

namespace fbk {
using namespace wah;

constexpr int CT_TPB = 256;
constexpr int CT_ITEMS = 4;
constexpr int CT = CT_TPB * CT_ITEMS;        // groups per tile
constexpr int CT_NW = CT_TPB / WAVE;
constexpr uint32_t CHUNKW = 32;
constexpr word_t SENT = 0xFFFFFFFFu;         // "no group" (no group value has bit 31)

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
    uint32_t ntiles;
    int op;
    word_t* out;
    uint32_t* outck;
    uint64_t* status;
    uint32_t* ctl;          // [1] words [3] groups [4],[5] edge words
};

// Tile starts of a compressed operand from its chunk index: the thread of
// chunk j (code words [32j, 32j+32), groups [ck[j], ck[j+1])) resolves the
// tiles whose first group falls in its span by walking those 32 words.
// ts[nt] = {n, groups} closes the table.
__device__ __forceinline__ void tile_starts(const word_t* __restrict__ w, const uint32_t* __restrict__ ck,
                                            uint32_t n, uint32_t nt, uint2* __restrict__ ts, uint32_t j) {
    const uint32_t nc = (n + CHUNKW - 1) / CHUNKW;
    if (j > nc) return;
    if (j == nc) { ts[nt] = make_uint2(n, ck[nc]); return; }
    const uint32_t c0 = ck[j], c1 = ck[j + 1];
    uint32_t t = (c0 + CT - 1) / CT;                      // first tile start >= c0
    if ((uint64_t)t * CT >= c1 || t >= nt) return;
    const uint32_t i0 = j * CHUNKW, e = min(i0 + CHUNKW, n);
    word_t x[CHUNKW];
    if (e == i0 + CHUNKW && (reinterpret_cast<uintptr_t>(w) & 15u) == 0) {
        const uint4* w4 = reinterpret_cast<const uint4*>(w + i0);
        #pragma unroll
        for (int q = 0; q < (int)CHUNKW / 4; ++q) {
            const uint4 u = w4[q];
            x[4 * q] = u.x; x[4 * q + 1] = u.y; x[4 * q + 2] = u.z; x[4 * q + 3] = u.w;
        }
    } else {
        #pragma unroll
        for (int q = 0; q < (int)CHUNKW; ++q) x[q] = i0 + (uint32_t)q < e ? w[i0 + q] : 0u;
    }
    uint32_t pos = c0;
    #pragma unroll
    for (int q = 0; q < (int)CHUNKW; ++q) {
        if (i0 + (uint32_t)q >= e || t >= nt) break;
        const uint32_t l = word_len(x[q]);
        while (t < nt && (uint64_t)t * CT < (uint64_t)pos + l) {
            ts[t] = make_uint2(i0 + (uint32_t)q, pos);
            ++t;
        }
        pos += l;
    }
}

// One launch before the tile pass: blockIdx.y = 0 / 1 builds the tile
// starts of compressed operand a / b, 2 clears the look-back status words
// and the control words.
__global__ __launch_bounds__(256) void k_canon_prep(CanonArgs p, uint32_t nt) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.y == 2) {
        if (j < 8u) p.ctl[j] = 0u;
        for (uint32_t i = j; i < nt; i += gridDim.x * blockDim.x) p.status[i] = 0ull;
        return;
    }
    const Operand& o = blockIdx.y == 0 ? p.a : p.b;
    if (!o.w || o.literal || o.n == 0) return;
    tile_starts(o.w, o.ck, o.n, nt, const_cast<uint2*>(o.ts), j);
}

__device__ __forceinline__ uint32_t cls3(word_t v) { return v == 0 ? 1u : (v == ALLONES ? 2u : 3u); }

// Inclusive max over the lanes of a wavefront (same DPP pattern as the sum).
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v) {
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
    return v;
}

// Exclusive block-wide max (values >= 0); sh needs NT/64 + 1 words.
template <int NT = CT_TPB>
__device__ __forceinline__ uint32_t block_excl_max(uint32_t v, uint32_t* sh) {
    const uint32_t w = threadIdx.x / WAVE;
    const uint32_t inc = wave_incl_max(v);
    if (lane_id() == WAVE - 1) sh[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t i = 0; i < w; ++i) before = max(before, sh[i]);
    uint32_t ex = (uint32_t)__shfl_up((int)inc, 1, 64);
    if (lane_id() == 0) ex = 0;
    __syncthreads();
    return max(before, ex);
}

// An operand's words for one tile, loaded before any is used so that both
// operands' loads are in flight together.  Literal operand: x[q] holds the
// tile's values (q < CT_ITEMS); compressed: x[r] the covering code words
// i0 + r*CT_TPB + t (r < CT_IT), p0 the first group of word i0.  Thread 0
// also loads the group values before and after the tile (e0 for a
// compressed operand only when word i0 starts at the tile).
constexpr int CT_IT = (CT + 1 + CT_TPB - 1) / CT_TPB;

struct Staged {
    uint32_t i0 = 0, p0 = 0, nw = 0, inx = 0;
    word_t x[CT_IT];
    word_t e0 = SENT, e1 = SENT;
};

__device__ __forceinline__ void stage_starts(const Operand& o, uint32_t tile, Staged& st) {
    if (o.literal) return;
    const uint2 a = o.ts[tile];
    st.i0 = a.x; st.p0 = a.y;
    st.inx = o.ts[tile + 1].x;
}

__device__ __forceinline__ void stage_words(const Operand& o, uint32_t g0, uint32_t cnt, uint32_t groups,
                                            Staged& st) {
    const int tid = threadIdx.x;
    const uint32_t g1 = g0 + cnt;
    if (o.literal) {
        #pragma unroll
        for (int q = 0; q < CT_ITEMS; ++q) {
            const uint32_t k = (uint32_t)q * CT_TPB + tid;
            st.x[q] = k < cnt ? o.w[g0 + k] : 0u;
        }
        if (tid == 0) {
            st.e0 = g0 > 0 ? o.w[g0 - 1] : SENT;
            st.e1 = g1 < groups ? o.w[g1] : SENT;
        }
        return;
    }
    // words covering the tile: i0 .. the word holding group g1 - 1
    st.nw = min(st.inx, o.n - 1u) + 1u - st.i0;          // <= CT + 1
    #pragma unroll
    for (int r = 0; r < CT_IT; ++r) {
        const uint32_t k = (uint32_t)r * CT_TPB + tid;
        st.x[r] = k < st.nw ? o.w[st.i0 + k] : 0u;
    }
    if (tid == 0) {
        st.e0 = (g0 > 0 && st.p0 == g0) ? group_value(o.w[st.i0 - 1u]) : SENT;
        st.e1 = g1 < groups ? group_value(o.w[st.inx]) : SENT;
    }
}

// Expands a staged operand over the tile's groups [g0, g0+cnt) into vals,
// combining with the values there by `op` when `combine` is set, and
// returns (to thread 0) the operand's values of groups g0-1 and g0+cnt
// (SENT where outside the bitvector).  Ends with a barrier.
__device__ void expand(const Operand& o, const Staged& st, uint32_t g0, uint32_t cnt, word_t* vals, word_t* sw,
                       uint32_t* sp, uint32_t* idx, uint32_t* shs, bool combine, int op, word_t& vprev,
                       word_t& vnext) {
    const int tid = threadIdx.x;
    if (o.literal) {
        #pragma unroll
        for (int q = 0; q < CT_ITEMS; ++q) {
            const uint32_t k = (uint32_t)q * CT_TPB + tid;
            if (k < cnt) vals[k] = combine ? apply_op(vals[k], st.x[q], op) : st.x[q];
        }
        if (tid == 0) { vprev = st.e0; vnext = st.e1; }
        __syncthreads();
        return;
    }
    const uint32_t nw = st.nw, p0 = st.p0;
    #pragma unroll
    for (int r = 0; r < CT_IT; ++r) {
        const uint32_t k = (uint32_t)r * CT_TPB + tid;
        if (k < nw) sw[k] = st.x[r];
    }
    for (uint32_t k = tid; k < cnt; k += CT_TPB) idx[k] = 0u;
    if (tid == 0) { vprev = st.e0; vnext = st.e1; }
    __syncthreads();
    if (tid == 0 && g0 > 0 && p0 < g0) vprev = group_value(sw[0]);
    // start groups of the staged words (thread-contiguous runs of CT_IT)
    uint32_t len[CT_IT];
    uint32_t s = 0;
    #pragma unroll
    for (int k = 0; k < CT_IT; ++k) {
        const uint32_t j = tid * CT_IT + k;
        len[k] = j < nw ? word_len(sw[j]) : 0u;
        s += len[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_sum<CT_TPB>(s, shs, &tot) + p0;
    #pragma unroll
    for (int k = 0; k < CT_IT; ++k) {
        const uint32_t j = tid * CT_IT + k;
        if (j <= nw) sp[j] = ex;
        ex += len[k];
    }
    __syncthreads();
    // mark each word at its first group in the tile (word 0 at group 0)
    for (uint32_t j = tid; j < nw; j += CT_TPB) {
        const uint32_t f = sp[j] > g0 ? sp[j] - g0 : 0u;
        if (f < cnt) idx[f] = j;
    }
    __syncthreads();
    // covering word of every group: inclusive max-scan of the marks
    {
        const uint32_t k0 = tid * CT_ITEMS;
        uint32_t m[CT_ITEMS], run = 0;
        #pragma unroll
        for (int i = 0; i < CT_ITEMS; ++i) {
            m[i] = k0 + i < cnt ? idx[k0 + i] : 0u;
            run = max(run, m[i]);
        }
        uint32_t e = block_excl_max(run, shs);
        #pragma unroll
        for (int i = 0; i < CT_ITEMS; ++i) {
            e = max(e, m[i]);
            if (k0 + i < cnt) idx[k0 + i] = e;
        }
    }
    __syncthreads();
    #pragma unroll
    for (int q = 0; q < CT_ITEMS; ++q) {
        const uint32_t k = (uint32_t)q * CT_TPB + tid;
        if (k < cnt) {
            const word_t x = group_value(sw[idx[k]]);
            vals[k] = combine ? apply_op(vals[k], x, op) : x;
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(CT_TPB) void k_canon_tiles(CanonArgs p) {
    __shared__ word_t s_vals[CT + 2];       // [0] = group before the tile, [cnt+1] after
    __shared__ word_t s_sw[2][CT + 1];      // staged code words of each operand
    __shared__ uint32_t s_sp[CT + 2];       // their start groups
    __shared__ uint32_t s_idx[CT];          // covering word of each group
    __shared__ uint32_t s_shs[CT_NW + 1];
    __shared__ uint32_t s_cnt[CT_ITEMS * CT_NW], s_lh[CT_ITEMS * CT_NW];
    __shared__ uint32_t s_tile[4];

    const int tid = threadIdx.x;
    const int lane = (int)lane_id();
    const uint32_t w = (uint32_t)tid / WAVE;
    const uint64_t lt = (1ull << lane) - 1ull, le = lt | (1ull << lane);
    word_t* vals = s_vals + 1;
    {
        const uint32_t tile = blockIdx.x;
        const uint32_t g0 = tile * (uint32_t)CT;
        const uint32_t cnt = min((uint32_t)CT, p.groups - g0);

        // both operands' tile starts, then both operands' words in flight
        Staged sa, sb;
        stage_starts(p.a, tile, sa);
        if (p.b.w) stage_starts(p.b, tile, sb);
        stage_words(p.a, g0, cnt, p.groups, sa);
        if (p.b.w) stage_words(p.b, g0, cnt, p.groups, sb);
        word_t vpa = SENT, vna = SENT, vpb = SENT, vnb = SENT;
        expand(p.a, sa, g0, cnt, vals, s_sw[0], s_sp, s_idx, s_shs, false, p.op, vpa, vna);
        if (p.b.w) expand(p.b, sb, g0, cnt, vals, s_sw[1], s_sp, s_idx, s_shs, true, p.op, vpb, vnb);
        if (tid == 0) {
            word_t vp = vpa, vn = vna;
            if (p.b.w) {
                if (vp != SENT) vp = apply_op(vpa, vpb, p.op);
                if (vn != SENT) vn = apply_op(vna, vnb, p.op);
            }
            s_vals[0] = vp;
            vals[cnt] = vn;
        }
        __syncthreads();

        // heads of this thread's groups; ballots per (q, wave) segment
        uint64_t hb[CT_ITEMS];
        #pragma unroll
        for (int q = 0; q < CT_ITEMS; ++q) {
            const uint32_t k = (uint32_t)q * CT_TPB + tid;
            bool head = false;
            if (k < cnt) {
                const word_t v = vals[k];
                const uint32_t cl = cls3(v);
                const word_t pv = vals[(int)k - 1];
                head = (cl == 3u) || (g0 + k == 0) || (pv == SENT) || (cls3(pv) != cl);
            }
            hb[q] = __ballot(head);
            if (lane == 0) {
                s_cnt[q * CT_NW + w] = (uint32_t)__popcll(hb[q]);
                s_lh[q * CT_NW + w] = hb[q] ? g0 + (uint32_t)q * CT_TPB + w * WAVE + (63u - __clzll(hb[q])) + 1u : 0u;
            }
        }
        __syncthreads();
        uint32_t tc = 0, th = 0;
        #pragma unroll
        for (int i = 0; i < CT_ITEMS * CT_NW; ++i) { tc += s_cnt[i]; th = max(th, s_lh[i]); }
        if (w == 0) {
            uint32_t ec, eh;
            lookback_ch_wave(p.status, tile, tc, th, ec, eh);
            if (lane == 0) { s_tile[1] = ec; s_tile[2] = eh; }
        }
        __syncthreads();
        const uint32_t base_cnt = s_tile[1], base_lh = s_tile[2];

        // write pass: the word of group g goes to slot (heads <= g) - 1
        uint32_t seg_cnt = base_cnt, seg_lh = base_lh;   // over the segments before (q, w)
        #pragma unroll
        for (int q = 0; q < CT_ITEMS; ++q) {
            uint32_t c = seg_cnt, h = seg_lh;
            for (uint32_t i = 0; i < w; ++i) { c += s_cnt[q * CT_NW + i]; h = max(h, s_lh[q * CT_NW + i]); }
            const uint32_t k = (uint32_t)q * CT_TPB + tid;
            if (k < cnt) {
                const uint32_t g = g0 + k;
                const word_t v = vals[k];
                const uint32_t cl = cls3(v);
                const bool head = (hb[q] >> lane) & 1ull;
                const uint32_t heads_le = c + (uint32_t)__popcll(hb[q] & lt) + (head ? 1u : 0u);
                const uint64_t mine = hb[q] & le;
                const uint32_t lh1 = mine ? g0 + (uint32_t)q * CT_TPB + w * WAVE + (63u - __clzll(mine)) + 1u : h;
                bool emit = false;
                word_t word = v;
                uint32_t headpos = g;
                if (cl == 3u) {
                    emit = true;
                } else {
                    const word_t nv = vals[k + 1];
                    const bool end = (g + 1 == p.groups) || (nv == SENT) || (cls3(nv) != cl);
                    if (end) {
                        emit = true;
                        headpos = lh1 - 1;
                        const uint32_t len = g - headpos + 1;
                        if (len >= 2) word = (cl == 2u ? HEADER1 : HEADER0) | len;
                    }
                }
                if (emit) {
                    const uint32_t slot = heads_le - 1;
                    p.out[slot] = word;
                    if ((slot & (CHUNKW - 1)) == 0 && slot < p.groups) p.outck[slot / CHUNKW] = headpos;
                }
                if (g + 1 == p.groups) {
                    p.ctl[1] = heads_le;
                    p.ctl[3] = p.groups;
                    p.outck[(heads_le + CHUNKW - 1) / CHUNKW] = p.groups;
                }
            }
            for (uint32_t i = 0; i < CT_NW; ++i) { seg_cnt += s_cnt[q * CT_NW + i]; seg_lh = max(seg_lh, s_lh[q * CT_NW + i]); }
        }
    }
}

// Literal-only operands (one word per group: the c0 routine and compress()
// of a literal array).  The 16-byte chunks (four groups) of the bitvector
// are split evenly over the waves of a grid that is resident at once, so
// every wave streams the same number of bytes.  A wave takes its span in
// rounds of 64 chunks, one per lane, and loads up to LF_R rounds at once.
// A group heads an output word when it is a mixed literal or its class
// (all-0 / all-1 / mixed) differs from the group before; it ends one when it
// is a mixed literal or its class differs from the group after.  A round's
// classes are wave-wide lane masks (one compare per group and class, the
// neighbours by DPP wave shifts, the counts by scalar popcounts); a round of
// mixed literals only, where every group is its own word, takes a shorter
// path.  The words a round ends take consecutive slots; they are stored
// straight from registers when the round's words are its values and the
// first slot is 16-byte aligned, else through a per-wave LDS row.
// Operands whose spans fit one load batch per wave take one launch
// (k_lit_fused, below); larger ones two, with no waiting between blocks:
//   k_lit_count: (heads, last head + 1) of every wave's span, and of every
//                block;
//   k_lit_write: each block sums the totals of the blocks before it (and of
//                its own waves before each wave), reads its span again and
//                writes the words.
constexpr int LW_TPB = 256;
constexpr int LW_NW = LW_TPB / WAVE;
constexpr uint32_t LW_ROW = 4u * WAVE;       // words one round can end
constexpr int LF_R = 5;                      // rounds a wave loads at once
constexpr int LW_PRE = 8;                    // block totals a thread sums
constexpr uint32_t LW_MAXB = LW_TPB * LW_PRE;

struct LitSpan {
    uint32_t c0, c1, nr;                     // chunks [c0, c1) in nr rounds
};

// Wave k's span: n4 / nwaves chunks, one more for the first n4 % nwaves waves.
__device__ __forceinline__ LitSpan lit_span(uint32_t n4, uint32_t k, uint32_t nwaves) {
    const uint32_t per = n4 / nwaves, rem = n4 % nwaves;
    LitSpan s;
    s.c0 = k * per + min(k, rem);
    s.c1 = s.c0 + per + (k < rem ? 1u : 0u);
    s.nr = (s.c1 - s.c0 + WAVE - 1) / WAVE;
    return s;
}

__device__ __forceinline__ word_t comp(const uint4& x, int j) {
    return j == 0 ? x.x : (j == 1 ? x.y : (j == 2 ? x.z : x.w));
}

template <int OP>
__device__ __forceinline__ word_t lit_op1(word_t a, word_t b) {
    if constexpr (OP < 0) return a;
    else return apply_op(a, b, OP);
}

template <int OP>
__device__ __forceinline__ uint4 lit_op4(const uint4& a, const uint4& b) {
    if constexpr (OP < 0) return a;
    else return make_uint4(apply_op(a.x, b.x, OP), apply_op(a.y, b.y, OP), apply_op(a.z, b.z, OP),
                           apply_op(a.w, b.w, OP));
}

// The op's value of group g (0 <= g < groups), a uniform load.
template <int OP>
__device__ __forceinline__ word_t lit_group(const CanonArgs& p, uint32_t g) {
    return lit_op1<OP>(p.a.w[g], OP < 0 ? 0u : p.b.w[g]);
}

// Uniform words of a span, every load in bounds (clamped): the groups
// before and after the span, the partial last chunk of the bitvector.
struct LitEdges {
    word_t before, after;
    uint4 tail;
    uint32_t n4f;                            // whole chunks (>= 1)
    uint32_t cmax;                           // last chunk a lane reads whole
};

template <int OP>
__device__ __forceinline__ LitEdges lit_edges(const CanonArgs& p, const LitSpan& s) {
    LitEdges e;
    const uint32_t groups = p.groups, gl = groups - 1u;
    e.n4f = groups / 4u;
    e.cmax = max(min(s.c1, e.n4f), 1u) - 1u;
    e.before = lit_group<OP>(p, max(4u * s.c0, 1u) - 1u);
    e.after = lit_group<OP>(p, min(4u * s.c1, gl));
    e.tail.x = lit_group<OP>(p, min(4u * e.n4f, gl));
    e.tail.y = lit_group<OP>(p, min(4u * e.n4f + 1u, gl));
    e.tail.z = lit_group<OP>(p, min(4u * e.n4f + 2u, gl));
    e.tail.w = SENT;
    if (s.c0 == 0) e.before = SENT;
    if (4u * s.c1 >= groups) e.after = SENT;
    if (4u * e.n4f >= groups) e.tail.x = SENT;
    if (4u * e.n4f + 1u >= groups) e.tail.y = SENT;
    if (4u * e.n4f + 2u >= groups) e.tail.z = SENT;
    return e;
}

// The op's values of rounds rb .. rb + LF_R - 1 of the span, every load in
// flight at once (lanes past the span read its last whole chunk; `last`
// masks them off later).
template <int OP>
__device__ __forceinline__ void lit_load(const CanonArgs& p, const LitSpan& s, const LitEdges& e, uint32_t rb,
                                         uint4 (&v)[LF_R]) {
    const uint32_t lane = lane_id();
    const uint4* A = reinterpret_cast<const uint4*>(p.a.w);
    const uint4* B = reinterpret_cast<const uint4*>(p.b.w);
    uint4 y[LF_R];
    #pragma unroll
    for (int r = 0; r < LF_R; ++r) {
        const uint32_t ic = min(s.c0 + (rb + (uint32_t)r) * WAVE + lane, e.cmax);
        v[r] = A[ic];
        if constexpr (OP >= 0) y[r] = B[ic];
    }
    #pragma unroll
    for (int r = 0; r < LF_R; ++r) {
        v[r] = lit_op4<OP>(v[r], y[r]);
        if (s.c0 + (rb + (uint32_t)r) * WAVE + lane == e.n4f) v[r] = e.tail;
    }
}

// Class masks of one round (bit l: lane l, group 4(i0 + l) + j).  Groups on
// lanes past `last`, or past the bitvector (SENT), are not valid.  Two fills
// of equal value are the same class, so a group differs in class from its
// neighbour exactly when it is mixed or their values differ.  prev / aft:
// the groups before lane 0 and after lane `last`.
struct LfMasks {
    uint64_t M[4], H[4], E[4];     // mixed, head, end
};

__device__ __forceinline__ uint64_t lanes_upto(uint32_t last) {
    return last >= (uint32_t)WAVE - 1u ? ~0ull : ((2ull << last) - 1ull);
}

__device__ __forceinline__ void lf_masks(const uint4& v, word_t prev, word_t aft, uint32_t last, LfMasks& m) {
    const uint64_t own = lanes_upto(last);
    const word_t xp = (word_t)__builtin_amdgcn_update_dpp((int)prev, (int)v.w, 0x138, 0xf, 0xf, false); // wave_shr:1
    word_t xn = (word_t)__builtin_amdgcn_update_dpp((int)aft, (int)v.x, 0x130, 0xf, 0xf, false);        // wave_shl:1
    if (lane_id() == last) xn = aft;
    const word_t b[6] = {xp, v.x, v.y, v.z, v.w, xn};
    uint64_t D[5];
    #pragma unroll
    for (int k = 0; k < 5; ++k) D[k] = ballot64(b[k] != b[k + 1]);
    #pragma unroll
    for (int j = 0; j < 4; ++j) {
        const word_t x = b[j + 1];
        const uint64_t V = ballot64(x != SENT) & own;
        const uint64_t F = ballot64(x == 0u || x == ALLONES);
        m.M[j] = V & ~F;
        m.H[j] = m.M[j] | (V & D[j]);
        m.E[j] = m.M[j] | (V & D[j + 1]);
    }
}

__device__ __forceinline__ uint32_t lane_bit(uint64_t mask) { return (uint32_t)(mask >> lane_id()) & 1u; }

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Every group of lanes 0..last is a mixed literal (neither fill nor SENT),
// so each is its own word, head and end.  x is mixed iff x - 1 < ALLONES - 1.
__device__ __forceinline__ bool lit_all_mixed(const uint4& v, uint32_t last) {
    const bool mixed = (v.x - 1u < ALLONES - 1u) && (v.y - 1u < ALLONES - 1u) && (v.z - 1u < ALLONES - 1u) &&
                       (v.w - 1u < ALLONES - 1u);
    return (ballot64(!mixed) & lanes_upto(last)) == 0;
}

// Heads of a round and the last head + 1 (0 when none); uniform.
__device__ __forceinline__ void lit_round_heads(const uint4& v, word_t prev, word_t aft, uint32_t i0,
                                                uint32_t last, uint32_t& nh, uint32_t& lh1) {
    if (lit_all_mixed(v, last)) {
        nh = 4u * (last + 1u);
        lh1 = 4u * (i0 + last) + 4u;
        return;
    }
    LfMasks m;
    lf_masks(v, prev, aft, last, m);
    nh = 0; lh1 = 0;
    #pragma unroll
    for (int j = 0; j < 4; ++j) {
        nh += (uint32_t)__popcll(m.H[j]);
        if (m.H[j]) lh1 = max(lh1, 4u * (i0 + 63u - (uint32_t)__clzll(m.H[j])) + (uint32_t)j + 1u);
    }
}

// Writes the words that a round ends, c heads (h: last head + 1) before it;
// `row` is the wave's LDS row; `closes` marks the round holding the
// bitvector's last group.  Returns the round's heads and last head + 1.
__device__ __forceinline__ void lit_round_write(const CanonArgs& p, const uint4& v, word_t prev, word_t aft,
                                                uint32_t i0, uint32_t last, bool closes, uint32_t c, uint32_t h,
                                                word_t* row, uint32_t& nh, uint32_t& lh1) {
    const uint32_t lane = lane_id();
    const uint32_t gl = 4u * (i0 + lane);
    uint32_t ne, sfirst;
    bool direct = false;
    if (lit_all_mixed(v, last)) {
        // the words are the values, from slot c
        nh = ne = 4u * (last + 1u);
        lh1 = 4u * (i0 + last) + 4u;
        sfirst = c;
        direct = (c & 3u) == 0;
        if (lane <= last) {
            if (direct) *reinterpret_cast<uint4*>(p.out + c + 4u * lane) = v;
            else reinterpret_cast<uint4*>(row)[lane] = v;
            #pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t slot = c + 4u * lane + (uint32_t)j;
                if ((slot & (CHUNKW - 1)) == 0 && slot < p.groups) p.outck[slot / CHUNKW] = gl + (uint32_t)j;
            }
        }
    } else {
        LfMasks m;
        lf_masks(v, prev, aft, last, m);
        nh = 0; lh1 = 0; ne = 0;
        uint32_t at = 0;                      // the lane's first row: ends on lanes below
        uint64_t fill_ends = 0;
        #pragma unroll
        for (int j = 0; j < 4; ++j) {
            nh += (uint32_t)__popcll(m.H[j]);
            if (m.H[j]) lh1 = max(lh1, 4u * (i0 + 63u - (uint32_t)__clzll(m.H[j])) + (uint32_t)j + 1u);
            at += lanes_below(m.E[j]);
            ne += (uint32_t)__popcll(m.E[j]);
            fill_ends |= m.E[j] & ~m.M[j];
        }
        // the round's first word is the one open at its first group, or the
        // one that group heads
        sfirst = c - 1u + (uint32_t)(m.H[0] & 1ull);
        if (!fill_ends) {
            // every word the round ends is a mixed literal: the value itself
            #pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (lane_bit(m.E[j])) {
                    row[at] = comp(v, j);
                    const uint32_t slot = sfirst + at;
                    if ((slot & (CHUNKW - 1)) == 0 && slot < p.groups) p.outck[slot / CHUNKW] = gl + (uint32_t)j;
                    ++at;
                }
            }
        } else {
            // a fill's word needs its run's first group: the last head at or
            // before it, in this lane or (a max-scan) before it
            uint32_t hb = 0;
            #pragma unroll
            for (int j = 0; j < 4; ++j) hb |= lane_bit(m.H[j]) << j;
            const uint32_t lhl = hb ? gl + (32u - __clz(hb)) : 0u;
            uint32_t lex = (uint32_t)__shfl_up((int)wave_incl_max(lhl), 1, 64);
            if (lane == 0) lex = 0;
            const uint32_t hl = max(h, lex);
            #pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (lane_bit(m.E[j])) {
                    const uint32_t g = gl + (uint32_t)j;
                    const word_t x = comp(v, j);
                    word_t word = x;
                    uint32_t headpos = g;
                    if (!lane_bit(m.M[j])) {
                        const uint32_t mine = hb & ((2u << j) - 1u);
                        headpos = mine ? gl + (31u - __clz(mine)) : hl - 1u;
                        const uint32_t len = g - headpos + 1u;
                        if (len >= 2u) word = (x == ALLONES ? HEADER1 : HEADER0) | len;
                    }
                    if (at < LW_ROW) row[at] = word;
                    const uint32_t slot = sfirst + at;
                    if ((slot & (CHUNKW - 1)) == 0 && slot < p.groups) p.outck[slot / CHUNKW] = headpos;
                    ++at;
                }
            }
        }
    }
    if (closes && lane == 0 && sfirst + ne <= p.groups) {
        // the bitvector's last group ends the last word
        p.ctl[1] = sfirst + ne;
        p.ctl[3] = p.groups;
        p.outck[(sfirst + ne + CHUNKW - 1) / CHUNKW] = p.groups;
    }
    if (!direct) {
        __builtin_amdgcn_wave_barrier();
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t idx = (uint32_t)q * WAVE + lane;
            if (idx < ne && sfirst + idx < p.groups) p.out[sfirst + idx] = row[idx];
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// The group after round R (= rb + r) of the span: the next round's first
// group (in registers within a batch), else a uniform load, else the
// group after the span.
template <int OP>
__device__ __forceinline__ word_t lit_after(const CanonArgs& p, const LitSpan& s, const LitEdges& e,
                                            const uint4 (&v)[LF_R], int r, uint32_t R) {
    if (R + 1u >= s.nr) return e.after;
    const int rn = r + 1 < LF_R ? r + 1 : r;
    if (r + 1 < LF_R) return (word_t)__builtin_amdgcn_readlane((int)v[rn].x, 0);
    return lit_group<OP>(p, 4u * (s.c0 + (R + 1u) * WAVE));
}

// status[k] = (heads, last head + 1) of wave k's span; status[nwaves + b]
// the same over block b.
template <int OP>
__global__ __launch_bounds__(LW_TPB) void k_lit_count(CanonArgs p, uint32_t n4, uint32_t nwaves) {
    __shared__ uint32_t s_c[LW_NW], s_h[LW_NW];
    const uint32_t tid = threadIdx.x;
    // the wave index is uniform: scalar span arithmetic and lane masks
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid / WAVE));
    const uint32_t k = blockIdx.x * LW_NW + wv;
    const LitSpan s = lit_span(n4, k, nwaves);
    const LitEdges e = lit_edges<OP>(p, s);
    uint32_t hc = 0, lh = 0;
    word_t prev = e.before;
    for (uint32_t rb = 0; rb < s.nr; rb += LF_R) {
        uint4 v[LF_R];
        lit_load<OP>(p, s, e, rb, v);
        #pragma unroll
        for (int r = 0; r < LF_R; ++r) {
            const uint32_t R = rb + (uint32_t)r;
            if (R >= s.nr) break;
            const uint32_t i0 = s.c0 + R * WAVE;
            const uint32_t last = min((uint32_t)WAVE - 1u, s.c1 - 1u - i0);
            uint32_t nh, lh1;
            lit_round_heads(v[r], prev, lit_after<OP>(p, s, e, v, r, R), i0, last, nh, lh1);
            hc += nh;
            lh = max(lh, lh1);
            prev = (word_t)__builtin_amdgcn_readlane((int)v[r].w, WAVE - 1);
        }
    }
    if (lane_id() == 0) {
        p.status[k] = ((uint64_t)hc << 32) | lh;
        s_c[wv] = hc; s_h[wv] = lh;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t c = 0, h = 0;
        for (int w = 0; w < LW_NW; ++w) { c += s_c[w]; h = max(h, s_h[w]); }
        p.status[nwaves + blockIdx.x] = ((uint64_t)c << 32) | h;
    }
}

template <int OP>
__global__ __launch_bounds__(LW_TPB) void k_lit_write(CanonArgs p, uint32_t n4, uint32_t nwaves) {
    __shared__ __align__(16) word_t s_row[LW_NW][LW_ROW];
    __shared__ uint32_t s_pc[LW_NW], s_ph[LW_NW], s_oc[LW_NW], s_oh[LW_NW];
    const uint32_t tid = threadIdx.x, lane = lane_id();
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid / WAVE));
    const uint32_t k = blockIdx.x * LW_NW + wv;
    const LitSpan s = lit_span(n4, k, nwaves);
    const LitEdges e = lit_edges<OP>(p, s);
    // totals of the blocks before this one (grid <= LW_MAXB) and of this
    // block's waves, summed before the operands are loaded: they are read
    // before the operand loads queue up, and their registers are free again
    uint32_t pc = 0, ph = 0;
    {
        uint64_t t[LW_PRE];
        const uint32_t jmax = max(blockIdx.x, 1u) - 1u;
        #pragma unroll
        for (int u = 0; u < LW_PRE; ++u) {
            const uint32_t j = tid + (uint32_t)u * LW_TPB;
            t[u] = p.status[nwaves + min(j, jmax)];
            if (j >= blockIdx.x) t[u] = 0;
        }
        #pragma unroll
        for (int u = 0; u < LW_PRE; ++u) { pc += (uint32_t)(t[u] >> 32); ph = max(ph, (uint32_t)t[u]); }
    }
    const uint64_t own = p.status[blockIdx.x * LW_NW + min(tid, (uint32_t)LW_NW - 1u)];
    pc = wave_sum(pc);
    ph = wave_max(ph);
    if (lane == 0) { s_pc[wv] = pc; s_ph[wv] = ph; }
    if (tid < LW_NW) { s_oc[tid] = (uint32_t)(own >> 32); s_oh[tid] = (uint32_t)own; }
    __syncthreads();
    uint32_t c = 0, h = 0;                    // heads before the round, last head + 1
    for (int w = 0; w < LW_NW; ++w) { c += s_pc[w]; h = max(h, s_ph[w]); }
    for (uint32_t w = 0; w < wv; ++w) { c += s_oc[w]; h = max(h, s_oh[w]); }
    word_t prev = e.before;
    for (uint32_t rb = 0; rb < s.nr; rb += LF_R) {
        uint4 v[LF_R];
        lit_load<OP>(p, s, e, rb, v);
        #pragma unroll
        for (int r = 0; r < LF_R; ++r) {
            const uint32_t R = rb + (uint32_t)r;
            if (R >= s.nr) break;
            const uint32_t i0 = s.c0 + R * WAVE;
            const uint32_t last = min((uint32_t)WAVE - 1u, s.c1 - 1u - i0);
            const bool closes = s.c1 == n4 && R + 1u == s.nr;
            uint32_t nh, lh1;
            lit_round_write(p, v[r], prev, lit_after<OP>(p, s, e, v, r, R), i0, last, closes, c, h, s_row[wv],
                            nh, lh1);
            c += nh;
            h = max(h, lh1);
            prev = (word_t)__builtin_amdgcn_readlane((int)v[r].w, WAVE - 1);
        }
    }
}

// Blocks of the literal path: resident at once, at most LW_MAXB, and no
// more than one round per wave needs.
static uint32_t lit_grid(uint32_t groups) {
    static const uint32_t slots = [] {
        int dev = 0, cus = 0, b1 = 0, b2 = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, k_lit_count<AND>, LW_TPB, 0) != hipSuccess || b1 <= 0)
            b1 = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b2, k_lit_write<AND>, LW_TPB, 0) != hipSuccess || b2 <= 0)
            b2 = 1;
        return min((uint32_t)cus * (uint32_t)min(b1, b2), LW_MAXB);
    }();
    const uint32_t rounds = ((groups + 3u) / 4u + WAVE - 1) / WAVE;
    return max(1u, min(slots, (rounds + LW_NW - 1) / LW_NW));
}

template <int OP>
static void launch_lit_op(const CanonArgs& p, uint32_t grid, hipStream_t s) {
    const uint32_t n4 = (p.groups + 3u) / 4u, nwaves = grid * LW_NW;
    hipLaunchKernelGGL(k_lit_count<OP>, dim3(grid), dim3(LW_TPB), 0, s, p, n4, nwaves);
    hipLaunchKernelGGL(k_lit_write<OP>, dim3(grid), dim3(LW_TPB), 0, s, p, n4, nwaves);
}

// groups >= 4 (a whole chunk to clamp loads to)
static void launch_canon_lit(const CanonArgs& p, hipStream_t s) {
    const uint32_t grid = lit_grid(p.groups);
    switch (p.b.w ? p.op : -1) {
    case AND: launch_lit_op<AND>(p, grid, s); break;
    case OR: launch_lit_op<OR>(p, grid, s); break;
    case XOR: launch_lit_op<XOR>(p, grid, s); break;
    case MINUS: launch_lit_op<MINUS>(p, grid, s); break;
    default: launch_lit_op<-1>(p, grid, s); break;
    }
}

// One launch for literal operands whose spans fit one load batch (LF_R
// rounds per wave): each block publishes a tagged total and sums the
// totals before it (blocks before it in its segment of 64 blocks, and the
// totals of earlier segments, which the last block of each segment
// publishes), polling with one wave and a sleep between passes; a block
// waits only for blocks of lower index, which the dispatcher has started
// before it.  It then writes its words from registers, so the operands
// are read once.  A total is one 64-bit word, [63:56] launch tag, [55:28]
// heads, [27:0] last head + 1; the tags cycle through 1..255 and the
// tables are cleared when they wrap.
constexpr uint32_t LF_SEG = 64;
// one polling lane per block of a segment
static_assert(LF_SEG <= (uint32_t)WAVE && LW_MAXB % LF_SEG == 0, "a segment is at most one wave of blocks");
__device__ uint64_t lf_status[LW_MAXB];
__device__ uint64_t lf_segtot[LW_MAXB / LF_SEG];

__global__ void k_lf_clear() {
    for (uint32_t j = threadIdx.x; j < LW_MAXB; j += blockDim.x) lf_status[j] = 0;
    for (uint32_t j = threadIdx.x; j < LW_MAXB / LF_SEG; j += blockDim.x) lf_segtot[j] = 0;
}

__device__ __forceinline__ uint64_t lf_pack(uint32_t tag, uint32_t cnt, uint32_t lh1) {
    return ((uint64_t)tag << 56) | ((uint64_t)(cnt & 0xFFFFFFFu) << 28) | (uint64_t)(lh1 & 0xFFFFFFFu);
}

// Block prefix of the one-launch kernels: tid 0 publishes the block's
// tagged total (bc, bh); wave 0 sums the totals before the block (those of
// the blocks before it in its segment, and the segment totals that the
// last block of every earlier segment publishes once its segment is in),
// polling with a sleep between passes and re-reading only what is missing.
// A block waits only for blocks of lower index, which the dispatcher has
// started before it.  Returns the sum in s_pc[0] / s_ph[0] to every thread.
__device__ __forceinline__ void lf_block_prefix(uint32_t tag, uint32_t bc, uint32_t bh, uint32_t wv,
                                                uint32_t* s_pc, uint32_t* s_ph) {
    const uint32_t tid = threadIdx.x, lane = lane_id();
    if (tid == 0)
        __hip_atomic_store(&lf_status[blockIdx.x], lf_pack(tag, bc, bh), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (wv == 0) {
        // lane l waits for block 64 sg + l (l < q) and segment l (l < sg)
        const uint32_t sg = blockIdx.x / LF_SEG, q = blockIdx.x % LF_SEG;
        bool need_b = lane < q, need_s = lane < sg;
        bool published = q != LF_SEG - 1u;
        uint32_t xc = 0, xh = 0, yc = 0, yh = 0;
        for (;;) {
            if (need_b) {
                const uint64_t x = __hip_atomic_load(&lf_status[sg * LF_SEG + lane], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                if ((uint32_t)(x >> 56) == tag) {
                    xc = (uint32_t)(x >> 28) & 0xFFFFFFFu; xh = (uint32_t)x & 0xFFFFFFFu; need_b = false;
                }
            }
            if (need_s) {
                const uint64_t x = __hip_atomic_load(&lf_segtot[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((uint32_t)(x >> 56) == tag) {
                    yc = (uint32_t)(x >> 28) & 0xFFFFFFFu; yh = (uint32_t)x & 0xFFFFFFFu; need_s = false;
                }
            }
            const bool pend_b = ballot64(need_b) != 0, pend_s = ballot64(need_s) != 0;
            if (!published && !pend_b) {
                // the last block of a segment publishes the segment's total
                const uint32_t tc = wave_sum(xc) + bc, th = max(wave_max(xh), bh);
                if (lane == 0)
                    __hip_atomic_store(&lf_segtot[sg], lf_pack(tag, tc, th), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                published = true;
            }
            if (!pend_b && !pend_s) break;
            __builtin_amdgcn_s_sleep(10);
        }
        const uint32_t pc = wave_sum(xc) + wave_sum(yc), ph = max(wave_max(xh), wave_max(yh));
        if (lane == 0) { s_pc[0] = pc; s_ph[0] = ph; }
    }
    __syncthreads();
}

// Launch tags of the one-launch kernels (shared table, so one sequence):
// 1..255 per device, the table cleared on the stream when they wrap.
static uint32_t lf_next_tag(hipStream_t s) {
    static std::mutex mu;
    static uint32_t tags[64];
    std::lock_guard<std::mutex> lk(mu);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    const uint32_t tag = tags[dev] % 255u + 1u;
    if (tag == 1u && tags[dev] != 0u) hipLaunchKernelGGL(k_lf_clear, dim3(1), dim3(1024), 0, s);
    tags[dev] = tag;
    return tag;
}

template <int OP>
__global__ __launch_bounds__(LW_TPB) void k_lit_fused(CanonArgs p, uint32_t n4, uint32_t nwaves, uint32_t tag) {
    __shared__ __align__(16) word_t s_row[LW_NW][LW_ROW];
    __shared__ uint32_t s_c[LW_NW], s_h[LW_NW], s_pc[1], s_ph[1];
    const uint32_t tid = threadIdx.x, lane = lane_id();
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid / WAVE));
    const uint32_t k = blockIdx.x * LW_NW + wv;
    const LitSpan s = lit_span(n4, k, nwaves);
    const LitEdges e = lit_edges<OP>(p, s);
    uint4 v[LF_R];
    lit_load<OP>(p, s, e, 0, v);
    uint32_t hc = 0, lh = 0;
    {
        word_t prev = e.before;
        #pragma unroll
        for (int r = 0; r < LF_R; ++r) {
            const uint32_t R = (uint32_t)r;
            if (R >= s.nr) break;
            const uint32_t i0 = s.c0 + R * WAVE;
            const uint32_t last = min((uint32_t)WAVE - 1u, s.c1 - 1u - i0);
            uint32_t nh, lh1;
            lit_round_heads(v[r], prev, lit_after<OP>(p, s, e, v, r, R), i0, last, nh, lh1);
            hc += nh;
            lh = max(lh, lh1);
            prev = (word_t)__builtin_amdgcn_readlane((int)v[r].w, WAVE - 1);
        }
    }
    if (lane == 0) { s_c[wv] = hc; s_h[wv] = lh; }
    __syncthreads();
    uint32_t bc = 0, bh = 0;
    for (int w = 0; w < LW_NW; ++w) { bc += s_c[w]; bh = max(bh, s_h[w]); }
    lf_block_prefix(tag, bc, bh, wv, s_pc, s_ph);
    uint32_t c = s_pc[0], h = s_ph[0];
    for (uint32_t w = 0; w < wv; ++w) { c += s_c[w]; h = max(h, s_h[w]); }
    word_t prev = e.before;
    #pragma unroll
    for (int r = 0; r < LF_R; ++r) {
        const uint32_t R = (uint32_t)r;
        if (R >= s.nr) break;
        const uint32_t i0 = s.c0 + R * WAVE;
        const uint32_t last = min((uint32_t)WAVE - 1u, s.c1 - 1u - i0);
        const bool closes = s.c1 == n4 && R + 1u == s.nr;
        uint32_t nh, lh1;
        lit_round_write(p, v[r], prev, lit_after<OP>(p, s, e, v, r, R), i0, last, closes, c, h, s_row[wv], nh, lh1);
        c += nh;
        h = max(h, lh1);
        prev = (word_t)__builtin_amdgcn_readlane((int)v[r].w, WAVE - 1);
    }
}

// Grid of the fused launch, 0 when the operands do not fit one batch.
static uint32_t lf_grid(uint32_t groups) {
    const uint32_t grid = lit_grid(groups);
    const uint32_t n4 = (groups + 3u) / 4u;
    return (n4 + grid * LW_NW - 1) / (grid * LW_NW) <= (uint32_t)(LF_R * WAVE) ? grid : 0u;
}

static void launch_lit_fused(const CanonArgs& p, uint32_t grid, hipStream_t s) {
    const uint32_t tag = lf_next_tag(s);
    const uint32_t n4 = (p.groups + 3u) / 4u, nwaves = grid * LW_NW;
    switch (p.b.w ? p.op : -1) {
    case AND: hipLaunchKernelGGL(k_lit_fused<AND>, dim3(grid), dim3(LW_TPB), 0, s, p, n4, nwaves, tag); break;
    case OR: hipLaunchKernelGGL(k_lit_fused<OR>, dim3(grid), dim3(LW_TPB), 0, s, p, n4, nwaves, tag); break;
    case XOR: hipLaunchKernelGGL(k_lit_fused<XOR>, dim3(grid), dim3(LW_TPB), 0, s, p, n4, nwaves, tag); break;
    case MINUS: hipLaunchKernelGGL(k_lit_fused<MINUS>, dim3(grid), dim3(LW_TPB), 0, s, p, n4, nwaves, tag); break;
    default: hipLaunchKernelGGL(k_lit_fused<-1>, dim3(grid), dim3(LW_TPB), 0, s, p, n4, nwaves, tag); break;
    }
}

// ---------------------------------------------------------------------------
// One launch for canonical ops with a compressed operand (the d1 / d2 / c1
// routines).  Rounds of 256 groups; wave k takes rounds [kR, kR + R).  A
// compressed operand's values of a round come from the 256 code words
// starting at the word holding the round's first group (256 words cover at
// least 256 groups): their start groups by a wave prefix sum, a mark at the
// start of each word in a per-wave LDS row, a max-scan of the marks for the
// word holding each group.  The word holding the next round's first group
// falls out of the same window; the span's first one comes from a 64-way
// search of the chunk index.  Classification, block totals and writes are
// those of k_lit_fused, from the values held in registers.

// The code word of compressed operand o holding group g (< its groups) and
// that word's first group, by one wave (uniform results).
__device__ __forceinline__ void mix_find_cover(const Operand& o, uint32_t g, uint32_t& wi, uint32_t& ws) {
    const uint32_t lane = lane_id();
    const uint32_t nc = (o.n + CHUNKW - 1) / CHUNKW;
    uint32_t lo = 0, hi = nc;               // ck[lo] <= g < ck[hi]; ck[nc] = groups
    while (hi - lo > 1u) {
        const uint32_t step = (hi - lo + WAVE - 1) / WAVE;
        const uint32_t j = lo + lane * step;
        const uint64_t m = ballot64(j < hi && o.ck[min(j, nc)] <= g);   // a prefix of the lanes
        lo += (63u - (uint32_t)__clzll(m)) * step;
        hi = min(lo + step, hi);
    }
    const uint32_t i = lo * CHUNKW + lane;
    const bool in = lane < CHUNKW && i < o.n;
    const uint32_t len = in ? word_len(o.w[min(i, o.n - 1u)]) : 0u;
    const uint32_t start = o.ck[lo] + wave_incl_sum(len) - len;
    const uint64_t cm = ballot64(in && start <= g && g - start < len);
    const uint32_t L = (uint32_t)__ffsll((unsigned long long)cm) - 1u;
    wi = lo * CHUNKW + L;
    ws = (uint32_t)__builtin_amdgcn_readlane((int)start, (int)L);
}

// mix_find_cover for two compressed operands at once: the levels of both
// searches advance together, so their dependent loads overlap.
__device__ __forceinline__ void mix_find_cover2(const Operand& a, const Operand& b, uint32_t g, uint32_t& wia,
                                                uint32_t& wsa, uint32_t& wib, uint32_t& wsb) {
    const uint32_t lane = lane_id();
    const uint32_t nca = (a.n + CHUNKW - 1) / CHUNKW, ncb = (b.n + CHUNKW - 1) / CHUNKW;
    uint32_t loa = 0, hia = nca, lob = 0, hib = ncb;
    while (hia - loa > 1u || hib - lob > 1u) {
        const uint32_t sa = (hia - loa + WAVE - 1) / WAVE, sb = (hib - lob + WAVE - 1) / WAVE;
        const uint32_t ja = loa + lane * sa, jb = lob + lane * sb;
        const bool ta = hia - loa > 1u && ja < hia && a.ck[min(ja, nca)] <= g;
        const bool tb = hib - lob > 1u && jb < hib && b.ck[min(jb, ncb)] <= g;
        const uint64_t ma = ballot64(ta), mb = ballot64(tb);
        if (hia - loa > 1u) { loa += (63u - (uint32_t)__clzll(ma)) * sa; hia = min(loa + sa, hia); }
        if (hib - lob > 1u) { lob += (63u - (uint32_t)__clzll(mb)) * sb; hib = min(lob + sb, hib); }
    }
    const uint32_t ia = loa * CHUNKW + lane, ib = lob * CHUNKW + lane;
    const bool ina = lane < CHUNKW && ia < a.n, inb = lane < CHUNKW && ib < b.n;
    const uint32_t lena = ina ? word_len(a.w[min(ia, a.n - 1u)]) : 0u;
    const uint32_t lenb = inb ? word_len(b.w[min(ib, b.n - 1u)]) : 0u;
    const uint32_t sta = a.ck[loa] + wave_incl_sum(lena) - lena;
    const uint32_t stb = b.ck[lob] + wave_incl_sum(lenb) - lenb;
    const uint64_t cma = ballot64(ina && sta <= g && g - sta < lena);
    const uint64_t cmb = ballot64(inb && stb <= g && g - stb < lenb);
    const uint32_t La = (uint32_t)__ffsll((unsigned long long)cma) - 1u;
    const uint32_t Lb = (uint32_t)__ffsll((unsigned long long)cmb) - 1u;
    wia = loa * CHUNKW + La;
    wsa = (uint32_t)__builtin_amdgcn_readlane((int)sta, (int)La);
    wib = lob * CHUNKW + Lb;
    wsb = (uint32_t)__builtin_amdgcn_readlane((int)stb, (int)Lb);
}

// Values of groups g0 + 4 lane + j of compressed operand o, whose word wi
// (starting at group ws <= g0) holds g0; wi / ws move on to the word holding
// g0 + 256.  mk / ww: the wave's LDS rows.
// The 256 words from wi (clamped to the operand's last word), 4 per lane.
__device__ __forceinline__ uint4 mix_window(const Operand& o, uint32_t wi) {
    const uint32_t i = wi + 4u * lane_id(), nl = o.n - 1u;
    return make_uint4(o.w[min(i, nl)], o.w[min(i + 1u, nl)], o.w[min(i + 2u, nl)], o.w[min(i + 3u, nl)]);
}

__device__ __forceinline__ uint4 mix_expand_round(const Operand& o, uint32_t g0, uint32_t& wi, uint32_t& ws,
                                                  const uint4& wv4, uint32_t* mk, word_t* ww) {
    const uint32_t lane = lane_id();
    const word_t w[4] = {wv4.x, wv4.y, wv4.z, wv4.w};
    uint32_t len[4], sum = 0;
    #pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t i = wi + 4u * lane + (uint32_t)j;
        len[j] = i < o.n ? word_len(w[j]) : 0u;
        sum += len[j];
    }
    uint32_t st = ws + wave_incl_sum(sum) - sum;          // first group of the lane's word 0
    #pragma unroll
    for (int j = 0; j < 4; ++j) { ww[4u * lane + j] = w[j]; mk[4u * lane + j] = 0u; }
    __builtin_amdgcn_wave_barrier();
    const uint32_t gn = g0 + 256u;
    uint32_t cand = 0, cst = 0, cen = 0;                  // last word starting at or before gn
    #pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t k = 4u * lane + (uint32_t)j;
        if (len[j] && st > g0 && st - g0 < 256u) mk[st - g0] = k;
        if (len[j] && st <= gn) { cand = k + 1u; cst = st; cen = st + len[j]; }
        st += len[j];
    }
    const uint32_t kn = wave_max(cand);                   // >= 1: word 0 starts at or before g0
    const uint32_t wsn = (uint32_t)__builtin_amdgcn_readlane((int)cst, (int)((kn - 1u) / 4u));
    const uint32_t wen = (uint32_t)__builtin_amdgcn_readlane((int)cen, (int)((kn - 1u) / 4u));
    __builtin_amdgcn_wave_barrier();
    const uint32_t m0 = mk[4u * lane], m1 = max(m0, mk[4u * lane + 1]);
    const uint32_t m2 = max(m1, mk[4u * lane + 2]), m3 = max(m2, mk[4u * lane + 3]);
    uint32_t ex = (uint32_t)__shfl_up((int)wave_incl_max(m3), 1, 64);
    if (lane == 0) ex = 0;
    uint4 r;
    r.x = group_value(ww[max(ex, m0)]);
    r.y = group_value(ww[max(ex, m1)]);
    r.z = group_value(ww[max(ex, m2)]);
    r.w = group_value(ww[max(ex, m3)]);
    __builtin_amdgcn_wave_barrier();
    // the window reaches gn (its words cover [ws, >= g0 + 256)): the word
    // found holds gn, or ends exactly there and the next word does
    if (wen <= gn) { wi += kn; ws = wen; }
    else { wi += kn - 1u; ws = wsn; }
    return r;
}

// Values of groups g0 + 4 lane + j of a literal operand (clamped reads).
__device__ __forceinline__ uint4 mix_lit_round(const Operand& o, uint32_t g0, uint32_t groups) {
    const uint32_t g = g0 + 4u * lane_id(), gl = groups - 1u;
    return make_uint4(o.w[min(g, gl)], o.w[min(g + 1u, gl)], o.w[min(g + 2u, gl)], o.w[min(g + 3u, gl)]);
}

// Value of group g of an operand, given for a compressed one the word wi
// (starting at ws) that holds group g + 1.
__device__ __forceinline__ word_t mix_group_before(const Operand& o, uint32_t g, uint32_t wi, uint32_t ws) {
    if (o.literal) return o.w[g];
    return group_value(o.w[ws <= g ? wi : (wi ? wi - 1u : 0u)]);
}

template <int OP>
__global__ __launch_bounds__(LW_TPB) void k_mix_fused(CanonArgs p, uint32_t nrounds, uint32_t R, uint32_t tag) {
    __shared__ __align__(16) word_t s_row[LW_NW][LW_ROW];
    __shared__ uint32_t s_mk[LW_NW][LW_ROW];
    __shared__ uint32_t s_c[LW_NW], s_h[LW_NW], s_pc[1], s_ph[1];
    const uint32_t lane = lane_id();
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const uint32_t k = blockIdx.x * LW_NW + wv;
    const uint32_t groups = p.groups, n4 = (groups + 3u) / 4u;
    // the span: rounds [m0, m0 + R) as 4-group chunks
    const uint32_t m0 = min(k * R, nrounds);
    LitSpan s;
    s.c0 = min(m0 * WAVE, n4);
    s.c1 = min((m0 + R) * WAVE, n4);
    s.nr = (s.c1 - s.c0 + WAVE - 1) / WAVE;
    const uint32_t G0 = 4u * s.c0, G1 = 4u * s.c1;
    const bool two = OP >= 0;
    uint32_t wia = 0, wsa = 0, wib = 0, wsb = 0;
    if (s.nr && two && !p.a.literal && !p.b.literal) {
        mix_find_cover2(p.a, p.b, G0, wia, wsa, wib, wsb);
    } else {
        if (s.nr && !p.a.literal) mix_find_cover(p.a, G0, wia, wsa);
        if (two && s.nr && !p.b.literal) mix_find_cover(p.b, G0, wib, wsb);
    }
    word_t before = SENT;
    if (s.nr && G0 > 0) {
        const word_t x = mix_group_before(p.a, G0 - 1u, wia, wsa);
        before = two ? lit_op1<OP>(x, mix_group_before(p.b, G0 - 1u, wib, wsb)) : x;
    }

    // phase 1: the span's values
    uint4 v[LF_R];
    #pragma unroll
    for (int r = 0; r < LF_R; ++r) {
        if ((uint32_t)r >= s.nr) break;
        const uint32_t g0 = G0 + 256u * (uint32_t)r;
        // both operands' windows (or literal groups) in flight before either
        // is expanded
        const uint4 wa = p.a.literal ? mix_lit_round(p.a, g0, groups) : mix_window(p.a, wia);
        uint4 wb = make_uint4(0, 0, 0, 0);
        if (two) wb = p.b.literal ? mix_lit_round(p.b, g0, groups) : mix_window(p.b, wib);
        const uint4 va = p.a.literal ? wa : mix_expand_round(p.a, g0, wia, wsa, wa, s_mk[wv], s_row[wv]);
        uint4 x = va;
        if (two) {
            const uint4 vb = p.b.literal ? wb : mix_expand_round(p.b, g0, wib, wsb, wb, s_mk[wv], s_row[wv]);
            x = lit_op4<OP>(va, vb);
        }
        const uint32_t g = g0 + 4u * lane;
        if (g >= groups) x.x = SENT;
        if (g + 1u >= groups) x.y = SENT;
        if (g + 2u >= groups) x.z = SENT;
        if (g + 3u >= groups) x.w = SENT;
        v[r] = x;
    }
    // the group after the span: held by the words the expansion moved to
    word_t after = SENT;
    if (s.nr && G1 < groups) {
        const word_t x = p.a.literal ? p.a.w[G1] : group_value(p.a.w[min(wia, p.a.n - 1u)]);
        after = x;
        if (two) after = lit_op1<OP>(x, p.b.literal ? p.b.w[G1] : group_value(p.b.w[min(wib, p.b.n - 1u)]));
    }

    // phase 2: heads of the span
    uint32_t hc = 0, lh = 0;
    {
        word_t prev = before;
        #pragma unroll
        for (int r = 0; r < LF_R; ++r) {
            const uint32_t Rr = (uint32_t)r;
            if (Rr >= s.nr) break;
            const uint32_t i0 = s.c0 + Rr * WAVE;
            const uint32_t last = min((uint32_t)WAVE - 1u, s.c1 - 1u - i0);
            const int rn = r + 1 < LF_R ? r + 1 : r;
            const word_t aft = Rr + 1u < s.nr ? (word_t)__builtin_amdgcn_readlane((int)v[rn].x, 0) : after;
            uint32_t nh, lh1;
            lit_round_heads(v[r], prev, aft, i0, last, nh, lh1);
            hc += nh;
            lh = max(lh, lh1);
            prev = (word_t)__builtin_amdgcn_readlane((int)v[r].w, WAVE - 1);
        }
    }
    if (lane == 0) { s_c[wv] = hc; s_h[wv] = lh; }
    __syncthreads();
    uint32_t bc = 0, bh = 0;
    for (int w = 0; w < LW_NW; ++w) { bc += s_c[w]; bh = max(bh, s_h[w]); }

    // phase 3: the totals before this block
    lf_block_prefix(tag, bc, bh, wv, s_pc, s_ph);
    uint32_t c = s_pc[0], h = s_ph[0];
    for (uint32_t w = 0; w < wv; ++w) { c += s_c[w]; h = max(h, s_h[w]); }

    // phase 4: the words, from registers
    word_t prev = before;
    #pragma unroll
    for (int r = 0; r < LF_R; ++r) {
        const uint32_t Rr = (uint32_t)r;
        if (Rr >= s.nr) break;
        const uint32_t i0 = s.c0 + Rr * WAVE;
        const uint32_t last = min((uint32_t)WAVE - 1u, s.c1 - 1u - i0);
        const int rn = r + 1 < LF_R ? r + 1 : r;
        const word_t aft = Rr + 1u < s.nr ? (word_t)__builtin_amdgcn_readlane((int)v[rn].x, 0) : after;
        const bool closes = s.c1 == n4 && Rr + 1u == s.nr;
        uint32_t nh, lh1;
        lit_round_write(p, v[r], prev, aft, i0, last, closes, c, h, s_row[wv], nh, lh1);
        c += nh;
        h = max(h, lh1);
        prev = (word_t)__builtin_amdgcn_readlane((int)v[r].w, WAVE - 1);
    }
}

// Rounds per wave of the mixed one-launch kernel (0: the operands need more
// than LF_R rounds per resident wave) and its grid.
static bool mix_grid(uint32_t groups, uint32_t& R, uint32_t& grid) {
    static const uint32_t slots = [] {
        int dev = 0, cus = 0, b = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_mix_fused<AND>, LW_TPB, 0) != hipSuccess || b <= 0)
            b = 1;
        return min((uint32_t)cus * (uint32_t)b, LW_MAXB);
    }();
    const uint32_t nrounds = (groups + 255u) / 256u;
    R = (nrounds + slots * LW_NW - 1) / (slots * LW_NW);
    if (R > (uint32_t)LF_R) return false;
    R = max(R, 1u);
    const uint32_t nwaves = (nrounds + R - 1) / R;
    grid = (nwaves + LW_NW - 1) / LW_NW;
    return grid <= LW_MAXB;
}

static void launch_mix_fused(const CanonArgs& p, uint32_t R, uint32_t grid, hipStream_t s) {
    const uint32_t tag = lf_next_tag(s);
    const uint32_t nrounds = (p.groups + 255u) / 256u;
    switch (p.b.w ? p.op : -1) {
    case AND: hipLaunchKernelGGL(k_mix_fused<AND>, dim3(grid), dim3(LW_TPB), 0, s, p, nrounds, R, tag); break;
    case OR: hipLaunchKernelGGL(k_mix_fused<OR>, dim3(grid), dim3(LW_TPB), 0, s, p, nrounds, R, tag); break;
    case XOR: hipLaunchKernelGGL(k_mix_fused<XOR>, dim3(grid), dim3(LW_TPB), 0, s, p, nrounds, R, tag); break;
    case MINUS: hipLaunchKernelGGL(k_mix_fused<MINUS>, dim3(grid), dim3(LW_TPB), 0, s, p, nrounds, R, tag); break;
    default: hipLaunchKernelGGL(k_mix_fused<-1>, dim3(grid), dim3(LW_TPB), 0, s, p, nrounds, R, tag); break;
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
    p.groups = groups; p.ntiles = groups ? nt : 0u;
    p.op = op; p.out = out; p.outck = outck; p.status = status; p.ctl = ctl;
    // (16-byte loads: operands at 16-byte aligned addresses; borrowed views
    // of index bitmaps may start anywhere and take the tile pass)
    const bool al16 = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15u) == 0;
    if (alit && (!b || blit) && groups >= 4u && al16) {
        const uint32_t fg = lf_grid(groups);
        if (fg) launch_lit_fused(p, fg, s);
        else launch_canon_lit(p, s);
        return hipGetLastError();
    }
    uint32_t mr = 0, mg = 0;
    if (groups && mix_grid(groups, mr, mg)) {
        launch_mix_fused(p, mr, mg, s);
        return hipGetLastError();
    }
    const uint32_t nca = !alit ? (na + CHUNKW - 1) / CHUNKW + 1 : 0u;
    const uint32_t ncb = b && !blit ? (nb + CHUNKW - 1) / CHUNKW + 1 : 0u;
    const uint32_t gx = max(max(nca, ncb), min(nt, 4096u * 256u));
    hipLaunchKernelGGL(k_canon_prep, dim3((gx + 255) / 256, 3), dim3(256), 0, s, p, nt);
    if (p.ntiles) hipLaunchKernelGGL(k_canon_tiles, dim3(p.ntiles), dim3(CT_TPB), 0, s, p);
    return hipGetLastError();
}

uint32_t canon_tiles(uint32_t groups) { return groups ? (groups + CT - 1) / CT : 1u; }

uint32_t canon_status(uint32_t groups) {
    return max(canon_tiles(groups), lit_grid(groups) * (LW_NW + 1));
}

} // namespace fbk
