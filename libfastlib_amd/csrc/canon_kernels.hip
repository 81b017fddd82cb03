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

// Clears the look-back status words and the control words of a launch.
__global__ void k_zero_canon(uint64_t* status, uint32_t n, uint32_t* ctl) {
    const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 < 8u) ctl[i0] = 0u;
    for (uint32_t i = i0; i < n; i += gridDim.x * blockDim.x) status[i] = 0ull;
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

// Expands operand `o` over the tile's groups [g0, g0+cnt) into vals,
// combining with the values there by `op` when `combine` is set, and
// returns (to thread 0) the operand's values of groups g0-1 and g0+cnt
// (SENT where outside the bitvector).  Ends with a barrier.
__device__ void expand(const Operand& o, uint32_t tile, uint32_t g0, uint32_t cnt, uint32_t groups,
                       word_t* vals, word_t* sw, uint32_t* sp, uint32_t* idx, uint32_t* shs, bool combine,
                       int op, word_t& vprev, word_t& vnext) {
    const int tid = threadIdx.x;
    const uint32_t g1 = g0 + cnt;
    if (o.literal) {
        word_t x[CT_ITEMS];
        #pragma unroll
        for (int q = 0; q < CT_ITEMS; ++q) {
            const uint32_t k = (uint32_t)q * CT_TPB + tid;
            x[q] = k < cnt ? o.w[g0 + k] : 0u;
        }
        #pragma unroll
        for (int q = 0; q < CT_ITEMS; ++q) {
            const uint32_t k = (uint32_t)q * CT_TPB + tid;
            if (k < cnt) vals[k] = combine ? apply_op(vals[k], x[q], op) : x[q];
        }
        if (tid == 0) {
            vprev = g0 > 0 ? o.w[g0 - 1] : SENT;
            vnext = g1 < groups ? o.w[g1] : SENT;
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
    #pragma unroll
    for (int r = 0; r < IT; ++r) {
        const uint32_t k = (uint32_t)r * CT_TPB + tid;
        if (k < nw) sw[k] = o.w[i0 + k];
    }
    for (uint32_t k = tid; k < cnt; k += CT_TPB) idx[k] = 0u;
    if (tid == 0) {
        vprev = g0 == 0 ? SENT : (p0 < g0 ? 0u : group_value(o.w[i0 - 1]));
        vnext = g1 < groups ? group_value(o.w[o.ts[tile + 1].x]) : SENT;
    }
    __syncthreads();
    if (tid == 0 && g0 > 0 && p0 < g0) vprev = group_value(sw[0]);
    // start groups of the staged words (thread-contiguous runs of IT)
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
    __shared__ word_t s_sw[CT + 1];         // staged code words
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

        word_t vpa = SENT, vna = SENT, vpb = SENT, vnb = SENT;
        expand(p.a, tile, g0, cnt, p.groups, vals, s_sw, s_sp, s_idx, s_shs, false, p.op, vpa, vna);
        if (p.b.w) expand(p.b, tile, g0, cnt, p.groups, vals, s_sw, s_sp, s_idx, s_shs, true, p.op, vpb, vnb);
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
                    if ((slot & (CHUNKW - 1)) == 0) p.outck[slot / CHUNKW] = headpos;
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

// Literal-only operands (one word per group, the c0 routine and compress()
// of a literal array) take three launches instead of one tile pass with a
// look-back: with thousands of tiles resident, waiting on predecessors'
// status words costs more than reading the operands twice (the second
// read mostly hits the last-level cache).
//   k_lit_count: per tile, the number of heads and the last head + 1;
//   k_lit_scan:  one block turns them into exclusive prefixes in place;
//   k_lit_write: per tile again, every output word at its slot.
// Lane l of wave w owns, in slice q of a tile, the four consecutive groups
// g0 + q*LT_SLICE + (w*64 + l)*4 + j, read with one 16-byte load per
// operand.  Neighbours of a lane's first and last group come from the
// adjacent lanes by shuffles, across waves and slices from two small LDS
// tables.  Head counts are summed per (slice, wave) segment, whose order is
// the group order.
constexpr uint32_t LT_SLICE = CT_TPB * 4u;   // groups per slice

template <int S>
struct LitTile {
    uint4 v[S];
    uint32_t heads = 0, ends = 0;     // bit 4q+j: group j of slice q heads / ends a word
};

__device__ __forceinline__ word_t comp(const uint4& x, int j) {
    return j == 0 ? x.x : (j == 1 ? x.y : (j == 2 ? x.z : x.w));
}

// Four groups from 16-byte group i4 of a literal operand (whole or partial
// at the end of the bitvector: no read past groups).
__device__ __forceinline__ uint4 load4(const word_t* w, uint32_t i4, uint32_t groups) {
    const uint32_t g = i4 * 4u;
    if (g + 4u <= groups) return reinterpret_cast<const uint4*>(w)[i4];
    uint4 r = make_uint4(SENT, SENT, SENT, SENT);
    if (g < groups) r.x = w[g];
    if (g + 1u < groups) r.y = w[g + 1u];
    if (g + 2u < groups) r.z = w[g + 2u];
    return r;
}

__device__ __forceinline__ uint4 op4(const uint4& a, const uint4& b, int op, uint32_t g, uint32_t groups) {
    uint4 r;
    r.x = g < groups ? apply_op(a.x, b.x, op) : SENT;
    r.y = g + 1u < groups ? apply_op(a.y, b.y, op) : SENT;
    r.z = g + 2u < groups ? apply_op(a.z, b.z, op) : SENT;
    r.w = g + 3u < groups ? apply_op(a.w, b.w, op) : SENT;
    return r;
}

// Loads the op's values of the lane's groups and classifies them.  The LDS
// tables need S*CT_NW words each, s_edge 2 words.  Ends with the head count
// and the last head + 1 of every (slice, wave) segment in s_cnt / s_lh.
template <int S>
struct LitRaw {
    uint4 x[S], y[S];
};

// Issues the loads of the tile at g0 (tiles past the end load nothing).
template <int S>
__device__ __forceinline__ void lit_load(const CanonArgs& p, uint32_t g0, LitRaw<S>& r) {
    const uint32_t groups = p.groups;
    const bool two = p.b.w != nullptr;
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        const uint32_t i4 = (g0 + (uint32_t)q * LT_SLICE) / 4u + (uint32_t)threadIdx.x;
        r.x[q] = load4(p.a.w, i4, groups);
        r.y[q] = two ? load4(p.b.w, i4, groups) : make_uint4(0u, 0u, 0u, 0u);
    }
}

template <int S>
__device__ __forceinline__ void lit_wait(const LitRaw<S>& r) {
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        use_loaded(r.x[q].x); use_loaded(r.x[q].y); use_loaded(r.x[q].z); use_loaded(r.x[q].w);
        use_loaded(r.y[q].x); use_loaded(r.y[q].y); use_loaded(r.y[q].z); use_loaded(r.y[q].w);
    }
}

template <int S>
__device__ __forceinline__ void lit_classify(const CanonArgs& p, uint32_t g0, LitTile<S>& T,
                                             word_t (*s_first)[CT_NW], word_t (*s_last)[CT_NW], word_t* s_edge,
                                             uint32_t* s_cnt, uint32_t* s_lh, const LitRaw<S>& r) {
    const int tid = threadIdx.x;
    const int lane = (int)lane_id();
    const uint32_t w = (uint32_t)tid / WAVE;
    const uint32_t groups = p.groups;
    const bool two = p.b.w != nullptr;
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        const uint32_t g = g0 + (uint32_t)q * LT_SLICE + (uint32_t)tid * 4u;
        T.v[q] = two ? op4(r.x[q], r.y[q], p.op, g, groups) : r.x[q];
    }
    if (tid == 0) {
        const uint32_t TILE = LT_SLICE * S;
        word_t vp = SENT, vn = SENT;
        if (g0 > 0) { vp = p.a.w[g0 - 1]; if (two) vp = apply_op(vp, p.b.w[g0 - 1], p.op); }
        if (g0 + TILE < groups) { vn = p.a.w[g0 + TILE]; if (two) vn = apply_op(vn, p.b.w[g0 + TILE], p.op); }
        s_edge[0] = vp; s_edge[1] = vn;
    }
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        if (lane == 0) s_first[q][w] = T.v[q].x;
        if (lane == WAVE - 1) s_last[q][w] = T.v[q].w;
    }
    __syncthreads();
    #pragma unroll
    for (int q = 0; q < S; ++q) {
        const uint32_t gl = g0 + (uint32_t)q * LT_SLICE + (uint32_t)tid * 4u;   // the lane's first group
        word_t pv = (word_t)__shfl_up((int)T.v[q].w, 1, 64);
        word_t nv = (word_t)__shfl_down((int)T.v[q].x, 1, 64);
        if (lane == 0) pv = w > 0 ? s_last[q][w - 1] : (q > 0 ? s_last[q - 1][CT_NW - 1] : s_edge[0]);
        if (lane == WAVE - 1) nv = w + 1 < CT_NW ? s_first[q][w + 1] : (q + 1 < S ? s_first[q + 1][0] : s_edge[1]);
        uint32_t hc = 0, lh1 = 0;
        #pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t g = gl + (uint32_t)j;
            const word_t v = comp(T.v[q], j);
            const word_t before = j == 0 ? pv : comp(T.v[q], j - 1);
            const word_t after = j == 3 ? nv : comp(T.v[q], j + 1);
            bool head = false, end = false;
            if (g < groups) {
                const uint32_t cl = cls3(v);
                head = (cl == 3u) || (g == 0) || (before == SENT) || (cls3(before) != cl);
                end = (cl == 3u) || (g + 1 == groups) || (after == SENT) || (cls3(after) != cl);
            }
            T.heads |= (head ? 1u : 0u) << (4 * q + j);
            T.ends |= (end ? 1u : 0u) << (4 * q + j);
            if (head) { ++hc; lh1 = g + 1; }
        }
        const uint32_t sc = wave_sum(hc);
        const uint32_t sh = wave_max(lh1);
        if (lane == 0) { s_cnt[q * CT_NW + w] = sc; s_lh[q * CT_NW + w] = sh; }
    }
    __syncthreads();
}

// The count and write kernels take tiles in grid-stride order (there is
// no dependence between tiles here), the grid one wave of resident blocks.
template <int S>
__global__ __launch_bounds__(CT_TPB) void k_lit_count(CanonArgs p, uint32_t nt) {
    __shared__ word_t s_first[S][CT_NW], s_last[S][CT_NW];
    __shared__ uint32_t s_cnt[S * CT_NW], s_lh[S * CT_NW];
    __shared__ word_t s_edge[2];
    // the next tile's loads are in flight while this tile is classified
    LitRaw<S> cur, nxt;
    if (blockIdx.x < nt) lit_load<S>(p, blockIdx.x * LT_SLICE * S, cur);
    for (uint32_t tile = blockIdx.x; tile < nt; tile += gridDim.x) {
        const uint32_t g0 = tile * LT_SLICE * S;
        const uint32_t tn = tile + gridDim.x;
        if (tn < nt) lit_load<S>(p, tn * LT_SLICE * S, nxt);
        lit_wait<S>(cur);
        LitTile<S> T;
        lit_classify<S>(p, g0, T, s_first, s_last, s_edge, s_cnt, s_lh, cur);
        if (threadIdx.x == 0) {
            uint32_t tc = 0, th = 0;
            for (int i = 0; i < S * CT_NW; ++i) { tc += s_cnt[i]; th = max(th, s_lh[i]); }
            p.status[tile] = ((uint64_t)tc << 32) | th;
        }
        __syncthreads();
        cur = nxt;
    }
}

// Exclusive (sum of counts, max of last heads) over the nt tile entries of
// status, in place, by one block of 1024 threads: coalesced loads into
// LDS, contiguous runs per thread, one block scan.
constexpr int LS_TPB = 1024;
constexpr uint32_t LS_MAX = 8192;            // tiles per pass through LDS

__global__ __launch_bounds__(LS_TPB) void k_lit_scan(uint64_t* status, uint32_t nt) {
    __shared__ uint64_t sv[LS_MAX];
    __shared__ uint32_t sh[LS_TPB / WAVE + 1];
    __shared__ uint32_t carry[2];
    if (threadIdx.x == 0) { carry[0] = 0; carry[1] = 0; }
    for (uint32_t b0 = 0; b0 < nt; b0 += LS_MAX) {
        const uint32_t n = min(LS_MAX, nt - b0);
        for (uint32_t i = threadIdx.x; i < n; i += LS_TPB) sv[i] = status[b0 + i];
        __syncthreads();
        const uint32_t per = (n + LS_TPB - 1) / LS_TPB;
        const uint32_t i0 = threadIdx.x * per, i1 = min(i0 + per, n);
        uint32_t c = 0, h = 0;
        for (uint32_t i = i0; i < i1; ++i) { c += (uint32_t)(sv[i] >> 32); h = max(h, (uint32_t)sv[i]); }
        uint32_t tot;
        uint32_t ce = block_excl_sum<LS_TPB>(c, sh, &tot) + carry[0];
        uint32_t he = max(block_excl_max<LS_TPB>(h, sh), carry[1]);
        for (uint32_t i = i0; i < i1; ++i) {
            const uint64_t v = sv[i];
            sv[i] = ((uint64_t)ce << 32) | he;
            ce += (uint32_t)(v >> 32);
            he = max(he, (uint32_t)v);
        }
        __syncthreads();
        if (threadIdx.x == LS_TPB - 1 || (i0 < n && i1 == n)) { carry[0] = ce; carry[1] = he; }
        for (uint32_t i = threadIdx.x; i < n; i += LS_TPB) status[b0 + i] = sv[i];
        __syncthreads();
    }
}

template <int S>
__global__ __launch_bounds__(CT_TPB) void k_lit_write(CanonArgs p, uint32_t nt) {
    __shared__ word_t s_first[S][CT_NW], s_last[S][CT_NW];
    __shared__ uint32_t s_cnt[S * CT_NW], s_lh[S * CT_NW];
    __shared__ word_t s_edge[2];
    const int tid = threadIdx.x;
    const int lane = (int)lane_id();
    const uint32_t w = (uint32_t)tid / WAVE;
    const uint32_t groups = p.groups;
    for (uint32_t tile = blockIdx.x; tile < nt; tile += gridDim.x) {
        const uint32_t g0 = tile * LT_SLICE * S;
        const uint64_t pre = p.status[tile];
        LitRaw<S> raw;
        lit_load<S>(p, g0, raw);
        lit_wait<S>(raw);
        LitTile<S> T;
        lit_classify<S>(p, g0, T, s_first, s_last, s_edge, s_cnt, s_lh, raw);
        uint32_t seg_cnt = (uint32_t)(pre >> 32), seg_lh = (uint32_t)pre;
        #pragma unroll
        for (int q = 0; q < S; ++q) {
            uint32_t c = seg_cnt, h = seg_lh;
            for (uint32_t i = 0; i < w; ++i) { c += s_cnt[q * CT_NW + i]; h = max(h, s_lh[q * CT_NW + i]); }
            const uint32_t gl = g0 + (uint32_t)q * LT_SLICE + (uint32_t)tid * 4u;
            const uint32_t hbits = (T.heads >> (4 * q)) & 0xFu, ebits = (T.ends >> (4 * q)) & 0xFu;
            // heads and last head + 1 of the lanes before this one in the wave
            const uint32_t hc = (uint32_t)__popc(hbits);
            const uint32_t cincl = wave_incl_sum(hc);
            const uint32_t lhl = hbits ? gl + (31u - __clz(hbits)) + 1u : 0u;
            uint32_t lex = (uint32_t)__shfl_up((int)wave_incl_max(lhl), 1, 64);
            if (lane == 0) lex = 0;
            c += cincl - hc;
            h = max(h, lex);
            if (ebits) {
                #pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (!((ebits >> j) & 1u)) continue;
                    const uint32_t g = gl + (uint32_t)j;
                    const uint32_t mine = hbits & ((2u << j) - 1u);
                    const uint32_t heads_le = c + (uint32_t)__popc(mine);
                    const uint32_t lh1 = mine ? gl + (31u - __clz(mine)) + 1u : h;
                    const word_t x = comp(T.v[q], j);
                    word_t word = x;
                    uint32_t headpos = g;
                    if (cls3(x) != 3u) {
                        headpos = lh1 - 1;
                        const uint32_t len = g - headpos + 1;
                        if (len >= 2) word = (x == ALLONES ? HEADER1 : HEADER0) | len;
                    }
                    const uint32_t slot = heads_le - 1;
                    p.out[slot] = word;
                    if ((slot & (CHUNKW - 1)) == 0) p.outck[slot / CHUNKW] = headpos;
                    if (g + 1 == groups) {
                        p.ctl[1] = heads_le;
                        p.ctl[3] = groups;
                        p.outck[(heads_le + CHUNKW - 1) / CHUNKW] = groups;
                    }
                }
            }
            for (uint32_t i = 0; i < CT_NW; ++i) { seg_cnt += s_cnt[q * CT_NW + i]; seg_lh = max(seg_lh, s_lh[q * CT_NW + i]); }
        }
        __syncthreads();        // the LDS tables are reused by the next tile
    }
}

template <int S>
static void launch_canon_lit(const CanonArgs& p, uint32_t groups, hipStream_t s) {
    static const uint32_t slots = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        return 8u * (uint32_t)cus;          // 8 blocks of 4 waves per CU
    }();
    const uint32_t nt = (groups + LT_SLICE * S - 1) / (LT_SLICE * S);
    const uint32_t grid = min(nt, slots);
    hipLaunchKernelGGL((k_lit_count<S>), dim3(grid), dim3(CT_TPB), 0, s, p, nt);
    hipLaunchKernelGGL(k_lit_scan, dim3(1), dim3(LS_TPB), 0, s, p.status, nt);
    hipLaunchKernelGGL((k_lit_write<S>), dim3(grid), dim3(CT_TPB), 0, s, p, nt);
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
    if (alit && (!b || blit) && groups) {
        launch_canon_lit<2>(p, groups, s);      // 2048-group tiles
        return hipGetLastError();
    }
    if (!alit && na > 0) {
        const uint32_t nc = (na + CHUNKW - 1) / CHUNKW;
        hipLaunchKernelGGL(k_tile_starts, dim3(nc / 256 + 1), dim3(256), 0, s, a, cka, na, nt, ts);
    }
    if (b && !blit && nb > 0) {
        const uint32_t nc = (nb + CHUNKW - 1) / CHUNKW;
        hipLaunchKernelGGL(k_tile_starts, dim3(nc / 256 + 1), dim3(256), 0, s, b, ckb, nb, nt, ts + (nt + 1));
    }
    hipLaunchKernelGGL(k_zero_canon, dim3(min((nt + 255) / 256, 4096u)), dim3(256), 0, s, status, nt, ctl);
    if (p.ntiles) hipLaunchKernelGGL(k_canon_tiles, dim3(p.ntiles), dim3(CT_TPB), 0, s, p);
    return hipGetLastError();
}

uint32_t canon_tiles(uint32_t groups) { return groups ? (groups + CT - 1) / CT : 1u; }

} // namespace fbk
