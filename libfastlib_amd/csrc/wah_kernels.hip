// gfx950 kernels for FastBit WAH bitvectors (32-bit words, 31-bit groups).
//
//  k_chunk_len      group count of every 32-word chunk (input of the chunk
//                   position index "ck": start group of code word 32*j)
//  k_op_tiles       a OP b over two compressed streams: merge-path over the
//                   chunk index, resynchronisation at points where the
//                   step state is fixed by the position (find_sync),
//                   the *_c2 step walk per thread, run-summary block scan and
//                   single-pass look-back, then the word write
//  k_tok_tiles      compressed output from a literal-domain token source:
//                   compress() of an arbitrary word array and the c0 path
//                   (both operands literal-only)
//  k_cnt            popcount + 31*(1-fill counts)
//  k_decompress     code words -> one literal word per group
//  k_flip           complement in place
//  k_rids           ascending positions of the set bits
#include "dev_common.h"

namespace fbk {
using namespace wah;

constexpr int TPB = 256;
constexpr int CHUNK = 32;           // code words per chunk-index entry
constexpr int KTILE = 64;           // chunk-index items per op tile
constexpr int WIN = CHUNK * (KTILE + 4);

// Code word accessor: LDS window [base, base+cnt) with global fallback.
struct WinReader {
    const word_t* lds;
    const word_t* g;
    uint32_t base;
    uint32_t cnt;
    __device__ __forceinline__ word_t at(uint32_t i) const {
        const uint32_t k = i - base;
        return k < cnt ? lds[k] : g[i];
    }
    static __device__ __forceinline__ WinReader pick(bool c, const WinReader& a, const WinReader& b) {
        return WinReader{c ? a.lds : b.lds, c ? a.g : b.g, c ? a.base : b.base, c ? a.cnt : b.cnt};
    }
};

struct GlobReader {
    const word_t* g;
    __device__ __forceinline__ word_t at(uint32_t i) const { return g[i]; }
    static __device__ __forceinline__ GlobReader pick(bool c, const GlobReader& a, const GlobReader& b) {
        return GlobReader{c ? a.g : b.g};
    }
};

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_chunk_len(const word_t* __restrict__ w, uint32_t n,
                                                   uint32_t* __restrict__ len, uint32_t nchunks) {
    // one wave per chunk group of 64 chunks: lane j sums chunk (base + j)
    const uint32_t j = blockIdx.x * TPB + threadIdx.x;
    if (j >= nchunks) return;
    const uint32_t b = j * CHUNK;
    const uint32_t e = min(b + CHUNK, n);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += word_len(w[i]);
    len[j] = s;
}

// Exclusive scan helpers over uint32 arrays (three-kernel form).
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = TPB * SCAN_ITEMS;

__global__ __launch_bounds__(TPB) void k_scan_tile(const uint32_t* __restrict__ in, uint32_t* out,
                                                   uint32_t n, uint32_t* __restrict__ tile_sums) {
    __shared__ uint32_t sh[TPB / WAVE + 1];
    const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
    #pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint32_t i = base + k;
        v[k] = i < n ? in[i] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t ex = block_excl_sum<TPB>(s, sh, &total);
    #pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint32_t i = base + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0 && tile_sums) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(TPB) void k_scan_add(uint32_t* out, uint32_t n,
                                                  const uint32_t* __restrict__ tile_off) {
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i < n) out[i] += tile_off[i / SCAN_TILE];
}

// Writes ck[nchunks] = total (the sentinel) once the scan is done.
__global__ void k_ck_sentinel(uint32_t* ck, uint32_t nchunks, const uint32_t* len) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        ck[nchunks] = nchunks ? ck[nchunks - 1] + len[nchunks - 1] : 0u;
}

// ---------------------------------------------------------------------------
struct OpArgs {
    const word_t* a; const uint32_t* cka; uint32_t na; uint32_t nca;
    const word_t* b; const uint32_t* ckb; uint32_t nb; uint32_t ncb;
    uint32_t groups;
    int op;
    uint32_t mode;        // MODE_CANON / MODE_RAW bits of the step walk
    word_t* out;
    uint32_t* outck;
    uint64_t* status;
    uint32_t* total;      // [0] output words, [3] groups represented
};

// Position of merged chunk-index item d (A items before B items on ties).
__device__ uint32_t diag_pos(const OpArgs& p, uint32_t d) {
    if (d >= p.nca + p.ncb) return p.groups;
    uint32_t lo = d > p.ncb ? d - p.ncb : 0u;
    uint32_t hi = d < p.nca ? d : p.nca;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (p.cka[mid] <= p.ckb[d - mid - 1]) lo = mid + 1; else hi = mid;
    }
    const uint32_t ja = lo, jb = d - lo;
    const uint32_t pa = ja < p.nca ? p.cka[ja] : 0xFFFFFFFFu;
    const uint32_t pb = jb < p.ncb ? p.ckb[jb] : 0xFFFFFFFFu;
    return pa < pb ? pa : pb;
}

// 64-way searches by one wave (uniform results; a few dependent loads
// instead of one per halving).
// The largest j < n with a[j] <= key (0 when none or n <= 1; a ascending).
__device__ __forceinline__ uint32_t wave_last_le(const uint32_t* a, uint32_t n, uint32_t key) {
    const uint32_t lane = lane_id();
    uint32_t lo = 0, hi = n;                 // a[lo] <= key (or lo = 0) < a[hi]
    while (hi - lo > 1u) {
        const uint32_t step = (hi - lo + WAVE - 1) / WAVE;
        const uint32_t j = lo + lane * step;
        const uint64_t m = ballot64(lane > 0 && j < hi && a[min(j, n - 1u)] <= key);
        const uint32_t L = m ? 63u - (uint32_t)__clzll(m) : 0u;
        lo += L * step;
        hi = min(lo + step, hi);
    }
    return lo;
}

// diag_pos by one wave: the first ja in [lo, hi] with cka[ja] > ckb[d - ja - 1].
__device__ uint32_t wave_diag_pos(const OpArgs& p, uint32_t d) {
    if (d >= p.nca + p.ncb) return p.groups;
    const uint32_t lane = lane_id();
    uint32_t lo = d > p.ncb ? d - p.ncb : 0u;
    uint32_t hi = d < p.nca ? d : p.nca;
    while (lo < hi) {
        const uint32_t step = (hi - lo + WAVE - 1) / WAVE;
        const uint32_t mid = lo + lane * step;
        const bool in = mid < hi;
        const uint32_t mc = in ? mid : lo;
        const uint64_t m = ballot64(in && p.cka[mc] <= p.ckb[d - mc - 1u]);   // a prefix of the lanes
        const uint32_t cnt = (uint32_t)__popcll(m);
        if (cnt == 0) { hi = lo; break; }
        const uint32_t nlo = lo + (cnt - 1u) * step + 1u;
        hi = min(lo + cnt * step, hi);
        lo = nlo;
    }
    const uint32_t ja = lo, jb = d - lo;
    const uint32_t pa = ja < p.nca ? p.cka[ja] : 0xFFFFFFFFu;
    const uint32_t pb = jb < p.ncb ? p.ckb[jb] : 0xFFFFFFFFu;
    return pa < pb ? pa : pb;
}

// Emitter for the write pass: result words and the result chunk index.
struct GlobWriter {
    word_t* out;
    uint32_t* outck;
    __device__ __forceinline__ void operator()(uint32_t slot, word_t w, uint32_t head) {
        out[slot] = w;
        if ((slot & (CHUNK - 1)) == 0) outck[slot / CHUNK] = head;
    }
};

// First code word starting at or after position P: looked up in the LDS
// window (start positions pw[0..cnt], pw[cnt] = end of the window) or, past
// the window, found by walking the global words.
__device__ __forceinline__ void holding_word(const uint32_t* pw, uint32_t cnt, uint32_t base,
                                             const word_t* g, uint32_t n, uint32_t P,
                                             uint32_t& i, uint32_t& pos) {
    // the word holding group P: pw[j] <= P < pw[j + 1]
    if (P < pw[cnt]) {
        const uint32_t j = last_le(pw, cnt + 1, P);
        i = base + j; pos = pw[j];
    } else {
        i = base + cnt; pos = pw[cnt];
        while (i < n) {
            const uint32_t l = word_len(g[i]);
            if (pos + l > P) break;
            pos += l; ++i;
        }
    }
}

__global__ __launch_bounds__(TPB) void k_op_tiles(OpArgs p) {
    __shared__ word_t sA[WIN];
    __shared__ word_t sB[WIN];
    __shared__ uint32_t posA[WIN + 1];
    __shared__ uint32_t posB[WIN + 1];
    __shared__ uint32_t tc[TPB + 1], tia[TPB + 1], tib[TPB + 1], tpa[TPB + 1], tpb[TPB + 1];
    __shared__ Summary sbuf[TPB / WAVE];
    __shared__ uint32_t info[20];
    __shared__ uint32_t shs[TPB / WAVE + 1];
    __shared__ Summary s_pre;

    const int tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;

    // the tile's two ends: wave 0 searches the first, wave 1 the second
    if (tid < 2 * WAVE) {
        const uint32_t e = (uint32_t)tid / WAVE;
        const uint32_t P = wave_diag_pos(p, (tile + e) * KTILE);
        const uint32_t ca = wave_last_le(p.cka, p.nca, P), cb = wave_last_le(p.ckb, p.ncb, P);
        if ((tid & (WAVE - 1)) == 0) { info[12 + 3 * e] = P; info[13 + 3 * e] = ca; info[14 + 3 * e] = cb; }
    }
    __syncthreads();
    if (tid == 0) {
        const uint32_t P0 = info[12], P1 = info[15];
        // chunk windows: from the chunk holding P0 through one chunk past
        // the chunk holding P1
        uint32_t ca0 = info[13];
        uint32_t ca1 = info[16] + 2;
        if (ca1 > p.nca) ca1 = p.nca;
        uint32_t cb0 = info[14];
        uint32_t cb1 = info[17] + 2;
        if (cb1 > p.ncb) cb1 = p.ncb;
        uint32_t baseA = ca0 * CHUNK, cntA = min(ca1 * CHUNK, p.na) - baseA;
        uint32_t baseB = cb0 * CHUNK, cntB = min(cb1 * CHUNK, p.nb) - baseB;
        if (cntA > WIN) cntA = WIN;
        if (cntB > WIN) cntB = WIN;
        info[1] = P0; info[2] = P1;
        info[3] = baseA; info[4] = cntA; info[5] = baseB; info[6] = cntB;
        info[7] = p.cka[ca0]; info[8] = p.ckb[cb0];
    }
    __syncthreads();
    const uint32_t P0 = info[1], P1 = info[2];
    const uint32_t baseA = info[3], cntA = info[4], baseB = info[5], cntB = info[6];

    // load windows (coalesced) and their start positions
    for (uint32_t i = tid; i < cntA; i += TPB) sA[i] = p.a[baseA + i];
    for (uint32_t i = tid; i < cntB; i += TPB) sB[i] = p.b[baseB + i];
    __syncthreads();
    {
        constexpr int IT = (WIN + 1 + TPB - 1) / TPB;
        uint32_t la[IT];
        uint32_t s = 0;
        #pragma unroll
        for (int k = 0; k < IT; ++k) {
            const uint32_t i = tid * IT + k;
            la[k] = i < cntA ? word_len(sA[i]) : 0u;
            s += la[k];
        }
        uint32_t tot;
        uint32_t ex = block_excl_sum<TPB>(s, shs, &tot) + info[7];
        #pragma unroll
        for (int k = 0; k < IT; ++k) {
            const uint32_t i = tid * IT + k;
            if (i <= cntA) posA[i] = ex;
            ex += la[k];
        }
        s = 0;
        #pragma unroll
        for (int k = 0; k < IT; ++k) {
            const uint32_t i = tid * IT + k;
            la[k] = i < cntB ? word_len(sB[i]) : 0u;
            s += la[k];
        }
        ex = block_excl_sum<TPB>(s, shs, &tot) + info[8];
        #pragma unroll
        for (int k = 0; k < IT; ++k) {
            const uint32_t i = tid * IT + k;
            if (i <= cntB) posB[i] = ex;
            ex += la[k];
        }
    }
    __syncthreads();

    WinReader ra{sA, p.a, baseA, cntA};
    WinReader rb{sB, p.b, baseB, cntB};

    // tile synchronisation points S0 = sync(P0), S1 = sync(P1), with the
    // words holding them and those words' starts
    if (tid == 0) {
        uint32_t S[2], IA[2], IB[2], PA[2], PB[2];
        const uint32_t Pq[2] = {P0, P1};
        for (int q = 0; q < 2; ++q) {
            if (Pq[q] >= p.groups) {
                S[q] = p.groups; IA[q] = p.na; IB[q] = p.nb; PA[q] = p.groups; PB[q] = p.groups;
                continue;
            }
            uint32_t ia, pa, ib, pb;
            holding_word(posA, cntA, baseA, p.a, p.na, Pq[q], ia, pa);
            holding_word(posB, cntB, baseB, p.b, p.nb, Pq[q], ib, pb);
            S[q] = find_sync(&ra, &rb, ia, pa, ib, pb, p.groups, (p.mode & MODE_RAW) != 0);
            IA[q] = ia; IB[q] = ib; PA[q] = pa; PB[q] = pb;
        }
        info[9] = S[0]; info[10] = IA[0]; info[11] = IB[0];
        info[12] = S[1]; info[13] = IA[1]; info[14] = IB[1];
        info[15] = PA[0]; info[16] = PB[0]; info[17] = PA[1]; info[18] = PB[1];
    }
    __syncthreads();
    const uint32_t S0 = info[9], S1 = info[12];

    // per-thread split: merge path over the window's code words in [S0,S1)
    {
        const uint32_t ia0 = info[10] - baseA, ib0 = info[11] - baseB;
        uint32_t iaE = info[13] - baseA, ibE = info[14] - baseB;
        if (iaE > cntA) iaE = cntA;
        if (ibE > cntB) ibE = cntB;
        const uint32_t nA = iaE > ia0 ? iaE - ia0 : 0u;
        const uint32_t nB = ibE > ib0 ? ibE - ib0 : 0u;
        const uint32_t tot = nA + nB;
        const uint32_t E = (tot + TPB - 1) / TPB;
        const uint32_t d = min((uint32_t)tid * E, tot);
        uint32_t c = S0, ia = info[10], ib = info[11], pa = info[15], pb = info[16];
        if (d != 0 && S0 < S1) {
            uint32_t lo = d > nB ? d - nB : 0u, hi = d < nA ? d : nA;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (posA[ia0 + mid] <= posB[ib0 + d - mid - 1]) lo = mid + 1; else hi = mid;
            }
            uint32_t ja = ia0 + lo, jb = ib0 + (d - lo);
            uint32_t qa = posA[ja], qb = posB[jb];
            // the split position is the next word start in merged order; the
            // other operand's previous word holds it.  A split at or before
            // S0 keeps the tile's own first point (the search is monotone).
            const uint32_t q0 = qa < qb ? qa : qb;
            if (q0 > S0) {
                if (qa < qb) { jb -= 1; qb = posB[jb]; }
                else if (qb < qa) { ja -= 1; qa = posA[ja]; }
                ja += baseA; jb += baseB;
                c = find_sync(&ra, &rb, ja, qa, jb, qb, S1, (p.mode & MODE_RAW) != 0);
                ia = ja; ib = jb; pa = qa; pb = qb;
                if (c >= S1) { c = S1; ia = info[13]; ib = info[14]; pa = info[17]; pb = info[18]; }
            }
        }
        tc[tid] = c; tia[tid] = ia; tib[tid] = ib; tpa[tid] = pa; tpb[tid] = pb;
        if (tid == 0) {
            tc[TPB] = S1; tia[TPB] = info[13]; tib[TPB] = info[14]; tpa[TPB] = info[17]; tpb[TPB] = info[18];
        }
    }
    __syncthreads();
    const uint32_t c0 = tc[tid], c1 = tc[tid + 1];

    // count pass
    CountSink cs;
    cs.init();
    if (c0 < c1) walk(cs, &ra, &rb, tia[tid], tpa[tid], tib[tid], tpb[tid], c0, c1, p.op, p.mode);
    Summary agg;
    const Summary texcl = block_excl_summary2<TPB>(cs.s, sbuf, &agg);
    if (tid < WAVE) {
        const Summary pr = prefix_summary_seg(p.status, p.status + gridDim.x, tile, 1u, agg);
        if (tid == 0) s_pre = pr;
    }
    __syncthreads();
    const Summary pre = combine(s_pre, texcl);

    // write pass
    GlobWriter gw{p.out, p.outck};
    if (c0 < c1) {
        WriteSink<GlobWriter> ws;
        ws.init(&gw, pre, c0);
        walk(ws, &ra, &rb, tia[tid], tpa[tid], tib[tid], tpb[tid], c0, c1, p.op, p.mode);
        if (c1 == p.groups) {
            ws.finish();
            const uint32_t total = combine(pre, cs.s).count;
            p.total[0] = total;
            p.outck[(total + CHUNK - 1) / CHUNK] = p.groups;
        }
    }
}

// ---------------------------------------------------------------------------
// Literal-domain token source kernels.  Tile = TPB*TOK_ITEMS source items.
constexpr int TOK_ITEMS = 16;
constexpr int TOK_TILE = TPB * TOK_ITEMS;

struct TokArgs {
    const word_t* a;      // compress: the words; c0: operand a (literal)
    const word_t* b;      // c0: operand b (literal); null for compress
    uint32_t n;           // source items
    uint32_t groups;      // total groups represented
    int op;
    int mode;             // 0 compress words (fills kept as F), 1 literal op, 2 literal copy
    word_t* out;
    uint32_t* outck;
    uint64_t* status;
    uint32_t* total;
};

template <class S>
__device__ __forceinline__ void tok_item(S& sink, word_t w, int mode) {
    if (mode == 0 && w > ALLONES) emit_fill(sink, w >= HEADER1 ? 1 : 0, w & MAXCNT, 1u);
    else emit_lit(sink, w, 1u);
}

__global__ __launch_bounds__(TPB) void k_tok_tiles(TokArgs p) {
    __shared__ word_t sw[TOK_TILE];
    __shared__ Summary sbuf[TPB / WAVE];
    __shared__ uint32_t shs[TPB / WAVE + 1];
    __shared__ uint32_t info[2];
    __shared__ Summary s_pre;
    __shared__ uint32_t s_base_pos;
    const int tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * TOK_TILE;
    const uint32_t cnt = base < p.n ? min((uint32_t)TOK_TILE, p.n - base) : 0u;
    // coalesced load of the tile's words (op applied for the c0 mode)
    for (uint32_t i = tid; i < cnt; i += TPB) {
        word_t w = p.a[base + i];
        if (p.mode == 1) w = apply_op(w, p.b[base + i], p.op);
        sw[i] = w;
    }
    __syncthreads();
    // positions: group count of each item (mode 0 only; else one group each)
    const uint32_t i0 = tid * TOK_ITEMS;
    uint32_t mylen = 0;
    #pragma unroll
    for (int k = 0; k < TOK_ITEMS; ++k) {
        const uint32_t i = i0 + k;
        if (i < cnt) mylen += (p.mode == 0) ? word_len(sw[i]) : 1u;
    }
    CountSink cs;
    cs.init();
    #pragma unroll 4
    for (int k = 0; k < TOK_ITEMS; ++k) {
        const uint32_t i = i0 + k;
        if (i < cnt) tok_item(cs, sw[i], p.mode);
    }
    uint32_t ltot;
    const uint32_t lpos = block_excl_sum<TPB>(mylen, shs, &ltot);
    Summary agg;
    const Summary texcl = block_excl_summary2<TPB>(cs.s, sbuf, &agg);
    if (tid < WAVE) {
        const Summary pr = prefix_summary_seg(p.status, p.status + gridDim.x, tile, 1u, agg);
        if (tid == 0) { s_pre = pr; s_base_pos = (p.mode == 0) ? 0u : base; }
    }
    __syncthreads();
    const Summary pre = combine(s_pre, texcl);
    GlobWriter gw{p.out, p.outck};
    WriteSink<GlobWriter> ws;
    ws.init(&gw, pre, s_base_pos + lpos);
    #pragma unroll 4
    for (int k = 0; k < TOK_ITEMS; ++k) {
        const uint32_t i = i0 + k;
        if (i < cnt) tok_item(ws, sw[i], p.mode);
    }
    const bool last = (cnt > 0) && (base + cnt == p.n) && (i0 < cnt) && (i0 + TOK_ITEMS >= cnt);
    if (last) {
        ws.finish();
        const uint32_t total = combine(pre, cs.s).count;
        p.total[0] = total;
        p.total[3] = ws.pos;                 // groups represented
        p.outck[(total + CHUNK - 1) / CHUNK] = ws.pos;
    }
}

// ---------------------------------------------------------------------------
// cnt / do_cnt in two passes without atomics: every block reduces a
// strided share of the words to (ones, groups) partials, one block sums them.
constexpr uint32_t CNT_BLOCKS = 512;

__global__ __launch_bounds__(TPB) void k_cnt(const word_t* __restrict__ w, uint32_t n,
                                             uint32_t* __restrict__ part) {
    __shared__ uint32_t so[TPB / WAVE], sg[TPB / WAVE];
    uint32_t ones = 0, groups = 0;
    const uint32_t stride = gridDim.x * TPB;
    uint32_t i = blockIdx.x * TPB + threadIdx.x;
    // four independent loads in flight per thread
    for (; i + 3u * stride < n; i += 4u * stride) {
        word_t x[4];
        #pragma unroll
        for (int r = 0; r < 4; ++r) x[r] = w[i + (uint32_t)r * stride];
        #pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (x[r] < HEADER0) { ones += __popc(x[r]); groups += 1; }
            else { const uint32_t c = x[r] & MAXCNT; groups += c; ones += (x[r] >= HEADER1) ? c * 31u : 0u; }
        }
    }
    for (; i < n; i += stride) {
        const word_t x = w[i];
        if (x < HEADER0) { ones += __popc(x); groups += 1; }
        else { const uint32_t c = x & MAXCNT; groups += c; ones += (x >= HEADER1) ? c * 31u : 0u; }
    }
    ones = wave_sum(ones);
    groups = wave_sum(groups);
    const uint32_t wv = threadIdx.x / WAVE;
    if (lane_id() == 0) { so[wv] = ones; sg[wv] = groups; }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0, b = 0;
        for (int k = 0; k < TPB / WAVE; ++k) { a += so[k]; b += sg[k]; }
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = b;
    }
}

__global__ __launch_bounds__(TPB) void k_cnt_final(const uint32_t* __restrict__ part, uint32_t nb,
                                                   uint32_t* __restrict__ out) {
    __shared__ uint32_t so[TPB / WAVE], sg[TPB / WAVE];
    uint32_t a = 0, b = 0;
    for (uint32_t k = threadIdx.x; k < nb; k += TPB) { a += part[2 * k]; b += part[2 * k + 1]; }
    a = wave_sum(a);
    b = wave_sum(b);
    const uint32_t wv = threadIdx.x / WAVE;
    if (lane_id() == 0) { so[wv] = a; sg[wv] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t x = 0, y = 0;
        for (int k = 0; k < TPB / WAVE; ++k) { x += so[k]; y += sg[k]; }
        out[0] = x;
        out[1] = y;
    }
}

// ---------------------------------------------------------------------------
// decompress: literal word per group.  One block per DEC_TILE output groups;
// the block finds the code words covering its groups through the chunk
// index, stages them and their start positions in LDS, and every thread
// resolves its groups by binary search.  neg != 0 writes the complement
// (copy_comp, src/bitvector.cpp:544-567).
constexpr int DEC_TILE = 2048;

__global__ __launch_bounds__(TPB) void k_decompress(const word_t* __restrict__ w, const uint32_t* ck,
                                                    uint32_t n, uint32_t nchunks, uint32_t groups,
                                                    word_t* __restrict__ out, int neg) {
    __shared__ word_t sw[DEC_TILE + 1];
    __shared__ uint32_t sp[DEC_TILE + 2];
    __shared__ uint32_t shs[TPB / WAVE + 1];
    __shared__ uint32_t info[3];
    const uint32_t g0 = blockIdx.x * DEC_TILE;
    if (g0 >= groups) return;
    const uint32_t g1 = min(g0 + (uint32_t)DEC_TILE, groups);
    if (threadIdx.x == 0) {
        const uint32_t c = last_le(ck, nchunks, g0);
        uint32_t i = c * CHUNK, pos = ck[c];
        while (i < n) {
            const uint32_t l = word_len(w[i]);
            if (pos + l > g0) break;
            pos += l; ++i;
        }
        info[0] = i; info[1] = pos;
    }
    __syncthreads();
    const uint32_t i0 = info[0], p0 = info[1];
    // at most DEC_TILE+1 code words can overlap DEC_TILE groups
    const uint32_t cnt = min((uint32_t)DEC_TILE + 1u, n - i0);
    for (uint32_t k = threadIdx.x; k < cnt; k += TPB) sw[k] = w[i0 + k];
    __syncthreads();
    constexpr int IT = (DEC_TILE + 1 + TPB - 1) / TPB;
    uint32_t l[IT];
    uint32_t s = 0;
    #pragma unroll
    for (int k = 0; k < IT; ++k) {
        const uint32_t j = threadIdx.x * IT + k;
        l[k] = j < cnt ? word_len(sw[j]) : 0u;
        s += l[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_sum<TPB>(s, shs, &tot) + p0;
    #pragma unroll
    for (int k = 0; k < IT; ++k) {
        const uint32_t j = threadIdx.x * IT + k;
        if (j <= cnt) sp[j] = ex;
        ex += l[k];
    }
    __syncthreads();
    for (uint32_t g = g0 + threadIdx.x; g < g1; g += TPB) {
        const uint32_t j = last_le(sp, cnt, g);
        const word_t x = sw[j];
        word_t v = group_value(x);
        out[g] = neg ? (v ^ ALLONES) : v;
    }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_flip(word_t* w, uint32_t n) {
    for (uint32_t i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) {
        const word_t x = w[i];
        w[i] = x > ALLONES ? (x ^ FILLBIT) : (x ^ ALLONES);
    }
}

// Zero fill for scratch (graph-capture friendly replacement of memset).
__global__ __launch_bounds__(TPB) void k_zero_u32(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) p[i] = 0u;
}

__global__ __launch_bounds__(TPB) void k_zero_u64(uint64_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) p[i] = 0ull;
}

} // namespace fbk

// ---------------------------------------------------------------------------
// Set-bit positions from a literal word array: one thread per word writes
// the positions of its set bits at its exclusive popcount offset.
namespace fbk {
__global__ __launch_bounds__(TPB) void k_popc(const word_t* __restrict__ lit, uint32_t n,
                                              uint32_t* __restrict__ cnt) {
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i < n) cnt[i] = __popc(lit[i]);
}

__global__ __launch_bounds__(TPB) void k_rids_lit(const word_t* __restrict__ lit, uint32_t n,
                                                  const uint32_t* __restrict__ off,
                                                  uint32_t* __restrict__ out, uint32_t cap) {
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    word_t w = lit[i];
    uint32_t o = off[i];
    const uint32_t pos0 = i * 31u;
    while (w) {
        const uint32_t hb = 31u - __clz(w);          // highest set bit
        if (o < cap) out[o] = pos0 + (30u - hb);
        ++o;
        w &= ~(1u << hb);
    }
}

// ---------------------------------------------------------------------------
// Host launchers.
static inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3(min(cdiv(n, TPB), 4096u)), dim3(TPB), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_zero_u64(uint64_t* p, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u64, dim3(min(cdiv(n, TPB), 4096u)), dim3(TPB), 0, s, p, n);
    return hipGetLastError();
}

// Exclusive scan; tmp needs scan_tmp_words(n) words.
uint32_t scan_tmp_words(uint32_t n) {
    uint32_t t = 0;
    while (n > (uint32_t)SCAN_TILE) { n = cdiv(n, SCAN_TILE); t += 2 * n + 1; }
    return t + 1;
}

hipError_t launch_scan_excl(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp,
                            hipStream_t s) {
    if (!n) return hipSuccess;
    const uint32_t nt = cdiv(n, SCAN_TILE);
    if (nt == 1) {
        hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(TPB), 0, s, in, out, n, (uint32_t*)nullptr);
        return hipGetLastError();
    }
    uint32_t* sums = tmp;
    uint32_t* offs = tmp + nt;
    hipLaunchKernelGGL(k_scan_tile, dim3(nt), dim3(TPB), 0, s, in, out, n, sums);
    hipError_t e = launch_scan_excl(sums, offs, nt, tmp + 2 * nt + 1, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3(cdiv(n, TPB)), dim3(TPB), 0, s, out, n, offs);
    return hipGetLastError();
}

// Chunk index ck[0..nchunks] of a word array (nchunks = ceil(n/32)).
// tmp needs nchunks + scan_tmp_words(nchunks) words.
hipError_t launch_chunk_index(const word_t* w, uint32_t n, uint32_t* ck, uint32_t* tmp,
                              hipStream_t s) {
    const uint32_t nc = cdiv(n, CHUNK);
    if (nc == 0) {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(TPB), 0, s, ck, 1u);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_chunk_len, dim3(cdiv(nc, TPB)), dim3(TPB), 0, s, w, n, tmp, nc);
    hipError_t e = launch_scan_excl(tmp, ck, nc, tmp + nc, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ck_sentinel, dim3(1), dim3(1), 0, s, ck, nc, (const uint32_t*)tmp);
    return hipGetLastError();
}

uint32_t op_tiles(uint32_t na, uint32_t nb) {
    const uint32_t items = cdiv(na, CHUNK) + cdiv(nb, CHUNK);
    return items ? cdiv(items, KTILE) : 1u;
}

// Run-summary entries of an op launch of nt tiles (prefix_summary_seg: the
// tiles' aggregates, then their segments').
uint32_t op_status_words(uint32_t nt) { return nt + nt / 64u + 1u; }

// Clears the look-back status words and the 8 control words of a launch.
__global__ void k_zero_status_ctl(uint64_t* status, uint32_t n, uint32_t* ctl) {
    const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 < 8u) ctl[i0] = 0u;
    for (uint32_t i = i0; i < n; i += gridDim.x * blockDim.x) status[i] = 0ull;
}

hipError_t launch_op(const word_t* a, const uint32_t* cka, uint32_t na,
                     const word_t* b, const uint32_t* ckb, uint32_t nb,
                     uint32_t groups, int op, int mode, word_t* out, uint32_t* outck,
                     uint64_t* status, uint32_t* ctl, hipStream_t s) {
    OpArgs p;
    p.a = a; p.cka = cka; p.na = na; p.nca = cdiv(na, CHUNK);
    p.b = b; p.ckb = ckb; p.nb = nb; p.ncb = cdiv(nb, CHUNK);
    p.groups = groups; p.op = op; p.mode = (uint32_t)mode;
    p.out = out; p.outck = outck; p.status = status;
    p.total = ctl + 1;
    const uint32_t nt = op_tiles(na, nb);
    const uint32_t ns = op_status_words(nt);
    hipLaunchKernelGGL(k_zero_status_ctl, dim3(min(cdiv(ns, TPB), 4096u)), dim3(TPB), 0, s, status, ns, ctl);
    hipLaunchKernelGGL(k_op_tiles, dim3(nt), dim3(TPB), 0, s, p);
    return hipGetLastError();
}

// ctl[4] = out[0], ctl[5] = out[ctl[1]-1] (first and last result word).
__global__ void k_edge_words(const word_t* out, uint32_t* ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t n = ctl[1];
        ctl[4] = n ? out[0] : 0u;
        ctl[5] = n ? out[n - 1] : 0u;
    }
}

// ---------------------------------------------------------------------------
// *_c2 results from the group domain.  A canonical op kernel has written the
// canonical words of a OP b; they are the *_c2 routines' words except that
// a result group forming a run of its own (a literal 0 / all-ones word) is a
// count-1 fill where the routine's token there is a fill (c2_rep_at).  Wave
// w fixes result words [256w, 256w + 256), lane l the four from 256w + 4l:
// their first groups from the result's chunk index and a wave prefix sum;
// the chunks of A and B that hold the wave's groups by 64-way searches (both
// ends, both operands, at once); then each lane's candidate words by a
// binary search within those chunks, one chunk's words, and the walk.
constexpr uint32_t FIX_WORDS = 256;
constexpr int FIX_PER = (int)FIX_WORDS / WAVE;          // result words per lane

// The chunk j of an operand with ck[j] <= g < ck[j + 1] by one wave.
__device__ __forceinline__ uint32_t wave_chunk_of(const uint32_t* ck, uint32_t nc, uint32_t g) {
    return wave_last_le(ck, nc + 1u, g);
}

// The chunk j of chunks [c0, c1) with ck[j] <= g < ck[j + 1], by one lane.
__device__ __forceinline__ uint32_t lane_chunk_of(const uint32_t* ck, uint32_t c0, uint32_t c1, uint32_t g) {
    uint32_t lo = c0, hi = c1;                 // ck[lo] <= g < ck[hi]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ck[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// The code word holding group g, in chunk j (first group c), and its first
// group, by one lane: the chunk's words eight at a time (two 16-byte loads;
// clamped scalar loads past the last word or at an unaligned address).
__device__ __forceinline__ void word_in_chunk(const word_t* w, uint32_t n, uint32_t j, uint32_t c, uint32_t g,
                                              uint32_t& wi, uint32_t& ws) {
    uint32_t pos = c, i = j * CHUNK;
    wi = i; ws = c;
    for (int piece = 0; piece < CHUNK / 8; ++piece, i += 8u) {
        word_t x[8];
        if (i + 8u <= n && (reinterpret_cast<uintptr_t>(w + i) & 15u) == 0) {
            const uint4* w4 = reinterpret_cast<const uint4*>(w + i);
            const uint4 u = w4[0], v = w4[1];
            x[0] = u.x; x[1] = u.y; x[2] = u.z; x[3] = u.w; x[4] = v.x; x[5] = v.y; x[6] = v.z; x[7] = v.w;
        } else {
            #pragma unroll
            for (int q = 0; q < 8; ++q) x[q] = w[min(i + (uint32_t)q, n - 1u)];
        }
        bool found = false;
        #pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t l = i + (uint32_t)q < n ? word_len(x[q]) : 0u;
            const bool here = !found && l && pos + l > g;
            wi = here ? i + (uint32_t)q : wi;
            ws = here ? pos : ws;
            found = found || here;
            pos += l;
        }
        if (found) break;
    }
}

// The words of chunks [c0, c1) of an operand staged in a wave's LDS rows
// (sw: words, sp: their first groups, then the end), when they number at
// most FIX_WIN; returns the count staged (0: too many, read from HBM).
// Lane l takes words 8l .. 8l + 7; all 64 lanes call it.
constexpr uint32_t FIX_WIN = 512;

__device__ __forceinline__ uint32_t stage_chunks(const word_t* w, const uint32_t* ck, uint32_t n, uint32_t c0,
                                                 uint32_t c1, word_t* sw, uint32_t* sp) {
    const uint32_t base = c0 * CHUNK, cnt = min(c1 * CHUNK, n) - base;
    if (cnt > FIX_WIN) return 0;
    const uint32_t lane = lane_id();
    word_t x[8];
    uint32_t len[8], sum = 0;
    #pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t j = 8u * lane + (uint32_t)k;
        x[k] = j < cnt ? w[base + j] : 0u;
        len[k] = j < cnt ? word_len(x[k]) : 0u;
        sum += len[k];
    }
    const uint32_t incl = wave_incl_sum(sum);
    uint32_t pos = ck[c0] + incl - sum;
    #pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t j = 8u * lane + (uint32_t)k;
        if (j < cnt) { sw[j] = x[k]; sp[j] = pos; }
        pos += len[k];
    }
    if (lane == WAVE - 1) sp[cnt] = pos;
    __builtin_amdgcn_wave_barrier();
    return cnt;
}

// The staged word holding group g (sp[0] <= g < sp[cnt]) and its first group.
__device__ __forceinline__ void staged_word_of(const uint32_t* sp, uint32_t cnt, uint32_t base, uint32_t g,
                                               uint32_t& wi, uint32_t& ws) {
    uint32_t lo = 0, hi = cnt;                 // sp[lo] <= g < sp[hi]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sp[mid] <= g) lo = mid; else hi = mid;
    }
    wi = base + lo; ws = sp[lo];
}

__global__ __launch_bounds__(TPB) void k_c2_fix(const word_t* __restrict__ a, const uint32_t* __restrict__ cka,
                                                uint32_t na, const word_t* __restrict__ b,
                                                const uint32_t* __restrict__ ckb, uint32_t nb, word_t* out,
                                                const uint32_t* __restrict__ outck, const uint32_t* ctl, int op) {
    __shared__ word_t s_w[TPB / WAVE][2][FIX_WIN];
    __shared__ uint32_t s_p[TPB / WAVE][2][FIX_WIN + 1];
    const uint32_t n = ctl[1];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x / WAVE;
    const uint32_t wbase = (blockIdx.x * (TPB / WAVE) + wv) * FIX_WORDS;
    if (wbase >= n) return;                                   // uniform per wave
    const uint32_t i0 = wbase + (uint32_t)FIX_PER * lane;
    word_t w[FIX_PER];
    uint32_t sum = 0, cand = 0;
    #pragma unroll
    for (int k = 0; k < FIX_PER; ++k) {
        w[k] = i0 + k < n ? out[i0 + k] : 0u;
        const uint32_t l = i0 + k < n ? word_len(w[k]) : 0u;
        sum += l;
        if (i0 + k < n && (w[k] == 0u || w[k] == ALLONES)) cand |= 1u << k;
    }
    if (!ballot64(cand != 0)) return;
    const uint32_t incl = wave_incl_sum(sum);
    const uint32_t P0 = outck[wbase / CHUNK];
    const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, WAVE - 1);
    // chunks of A and B that hold the wave's groups [P0, P0 + tot), staged
    // in LDS when they are few (dense results); else read from HBM
    const uint32_t nca = (na + CHUNK - 1) / CHUNK, ncb = (nb + CHUNK - 1) / CHUNK;
    const uint32_t glast = P0 + tot - 1u;
    const uint32_t ca0 = wave_chunk_of(cka, nca, P0), ca1 = min(wave_chunk_of(cka, nca, glast) + 1u, nca);
    const uint32_t cb0 = wave_chunk_of(ckb, ncb, P0), cb1 = min(wave_chunk_of(ckb, ncb, glast) + 1u, ncb);
    const uint32_t sa = stage_chunks(a, cka, na, ca0, ca1, s_w[wv][0], s_p[wv][0]);
    const uint32_t sb = stage_chunks(b, ckb, nb, cb0, cb1, s_w[wv][1], s_p[wv][1]);
    const WinReader ra{s_w[wv][0], a, ca0 * CHUNK, sa}, rb{s_w[wv][1], b, cb0 * CHUNK, sb};
    uint32_t pk[FIX_PER];                                     // first group of each word
    uint32_t pos = P0 + incl - sum;
    #pragma unroll
    for (int k = 0; k < FIX_PER; ++k) { pk[k] = pos; pos += word_len(w[k]); }
    // one candidate at a time (a select, not an indexed load, picks its
    // group and word), so the walk is not repeated per unrolled slot
    while (cand) {
        const uint32_t k = (uint32_t)__ffs(cand) - 1u;
        cand &= cand - 1u;
        uint32_t g = pk[0];
        word_t wk = w[0];
        #pragma unroll
        for (int q = 1; q < FIX_PER; ++q) {
            g = k == (uint32_t)q ? pk[q] : g;
            wk = k == (uint32_t)q ? w[q] : wk;
        }
        uint32_t ia, pa, ib, pb;
        if (sa) staged_word_of(s_p[wv][0], sa, ca0 * CHUNK, g, ia, pa);
        else word_in_chunk(a, na, lane_chunk_of(cka, ca0, ca1, g), cka[lane_chunk_of(cka, ca0, ca1, g)], g, ia, pa);
        if (sb) staged_word_of(s_p[wv][1], sb, cb0 * CHUNK, g, ib, pb);
        else word_in_chunk(b, nb, lane_chunk_of(ckb, cb0, cb1, g), ckb[lane_chunk_of(ckb, cb0, cb1, g)], g, ib, pb);
        if (c2_rep_at(&ra, &rb, ia, pa, ib, pb, g, op)) out[i0 + k] = (wk == ALLONES ? HEADER1 : HEADER0) | 1u;
    }
}

hipError_t launch_c2_fix(const word_t* a, const uint32_t* cka, uint32_t na, const word_t* b, const uint32_t* ckb,
                         uint32_t nb, word_t* out, const uint32_t* outck, const uint32_t* ctl, uint32_t maxwords,
                         int op, hipStream_t s) {
    const uint32_t waves = cdiv(maxwords, FIX_WORDS);
    if (!waves) return hipSuccess;
    hipLaunchKernelGGL(k_c2_fix, dim3(cdiv(waves, TPB / WAVE)), dim3(TPB), 0, s, a, cka, na, b, ckb, nb, out, outck,
                       ctl, op);
    return hipGetLastError();
}

hipError_t launch_edge_words(const word_t* out, uint32_t* ctl, hipStream_t s) {
    hipLaunchKernelGGL(k_edge_words, dim3(1), dim3(64), 0, s, out, ctl);
    return hipGetLastError();
}

uint32_t tok_tiles(uint32_t n) { return n ? cdiv(n, TOK_TILE) : 1u; }

// mode 0: compress `a` (n code words, `groups` groups); mode 1: canonical
// WAH of apply_op(a[i], b[i]) over n literal words; mode 2: canonical WAH of
// the literal words a[0..n).  status needs 2*tok_tiles(n) entries.
hipError_t launch_tok(const word_t* a, const word_t* b, uint32_t n, uint32_t groups, int op,
                      int mode, word_t* out, uint32_t* outck, uint64_t* status, uint32_t* ctl,
                      hipStream_t s) {
    TokArgs p;
    p.a = a; p.b = b; p.n = n; p.groups = groups; p.op = op; p.mode = mode;
    p.out = out; p.outck = outck; p.status = status; p.total = ctl + 1;
    const uint32_t nt = tok_tiles(n);
    hipLaunchKernelGGL(k_zero_u64, dim3(min(cdiv(2 * nt, TPB), 4096u)), dim3(TPB), 0, s, status,
                       2 * nt);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(TPB), 0, s, ctl, 8u);
    hipLaunchKernelGGL(k_tok_tiles, dim3(nt), dim3(TPB), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_cnt(const word_t* w, uint32_t n, uint32_t* out2, hipStream_t s) {
    uint32_t* part = out2 + 2;
    const uint32_t g = n ? min(cdiv(n, TPB * 4), CNT_BLOCKS) : 0u;
    if (g) hipLaunchKernelGGL(k_cnt, dim3(g), dim3(TPB), 0, s, w, n, part);
    hipLaunchKernelGGL(k_cnt_final, dim3(1), dim3(TPB), 0, s, part, g, out2);
    return hipGetLastError();
}

uint32_t cnt_scratch_words() { return 2u + 2u * CNT_BLOCKS; }

hipError_t launch_decompress(const word_t* w, const uint32_t* ck, uint32_t n, uint32_t groups,
                             word_t* out, int neg, hipStream_t s) {
    if (!groups) return hipSuccess;
    hipLaunchKernelGGL(k_decompress, dim3(cdiv(groups, DEC_TILE)), dim3(TPB), 0, s, w, ck, n,
                       cdiv(n, CHUNK), groups, out, neg);
    return hipGetLastError();
}

hipError_t launch_flip(word_t* w, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_flip, dim3(min(cdiv(n, TPB), 4096u)), dim3(TPB), 0, s, w, n);
    return hipGetLastError();
}

// rids of a literal array; tmp needs 2*n + scan_tmp_words(n) words.
hipError_t launch_rids_lit(const word_t* lit, uint32_t n, uint32_t* out, uint32_t cap,
                           uint32_t* tmp, hipStream_t s) {
    if (!n) return hipSuccess;
    uint32_t* cnt = tmp;
    uint32_t* off = tmp + n;
    hipLaunchKernelGGL(k_popc, dim3(cdiv(n, TPB)), dim3(TPB), 0, s, lit, n, cnt);
    hipError_t e = launch_scan_excl(cnt, off, n, tmp + 2 * n, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rids_lit, dim3(cdiv(n, TPB)), dim3(TPB), 0, s, lit, n,
                       (const uint32_t*)off, out, cap);
    return hipGetLastError();
}

} // namespace fbk
