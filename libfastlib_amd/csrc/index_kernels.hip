// gfx950 kernels for relic / bin indexes: build (keys, stable radix sort,
// per-bitmap canonical WAH encoding, concatenation, tile offset table) and
// range evaluation (LDS multiway OR with the edge-bin check, raw column scan).
//
// Index layout in HBM (the reference's in-memory form, src/irelic.cpp:644-667
// and src/ibin.cpp:5836-5860): bitmaps serialized one after another, bitmap
// b occupying words [off[b], off[b+1]) = its code words, then active.val when
// active.nbits > 0, then active.nbits.
//
// Tile offset table: for every tile t of TG 31-bit groups and every bitmap
// b, toff[t*nobs + b] = {index of the first code word of b whose span ends
// after group t*TG, start group of that word}.  A range query then lets one
// workgroup own one tile of the result and every lane walk the few words of
// one bitmap that overlap it.
#include "dev_common.h"
#include "index_dev.h"

namespace fbk {
using namespace wah;

constexpr int IX_TPB = 256;

// ---------------------------------------------------------------------------
// Keys.
template <class T>
__device__ __forceinline__ double as_double(T v) { return (double)v; }

// relic: index of the value in the sorted distinct values (lower bound).
template <class T>
__global__ __launch_bounds__(IX_TPB) void k_keys_relic(const T* __restrict__ col, uint32_t n,
                                                       const double* __restrict__ vals, uint32_t nv,
                                                       uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * IX_TPB + threadIdx.x;
    if (i >= n) return;
    const double x = as_double(col[i]);
    uint32_t lo = 0, hi = nv;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (vals[mid] < x) lo = mid + 1; else hi = mid;
    }
    keys[i] = lo;
}

// relic(col, card, ind) (src/irelic.cpp:160-211): the key is the value when
// it is below card, otherwise the row belongs to no bitmap.
__global__ __launch_bounds__(IX_TPB) void k_keys_card(const uint32_t* __restrict__ ind, uint32_t n,
                                                      uint32_t card, uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * IX_TPB + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = ind[i];
    keys[i] = v < card ? v : NOKEY;
}

// bin::locate(double) (src/ibin.cpp:673-738), restated exactly.
__device__ __forceinline__ uint32_t bin_locate(const double* bounds, uint32_t nobs, double val) {
    if (nobs == 0) return 0;
    if (!(val >= bounds[0])) return 0;
    if (!(val < bounds[nobs - 1])) return bounds[nobs - 1] < DBL_MAX_D ? nobs : nobs - 1;
    if (nobs >= 8) {
        uint32_t i0 = 0, i1 = nobs, it = nobs / 2;
        while (i0 < it) {
            if (val < bounds[it]) i1 = it; else i0 = it;
            it = (i0 + i1) / 2;
        }
        return i1;
    }
    for (uint32_t i = 1; i < nobs; ++i)
        if (val < bounds[i]) return i;
    return nobs;
}

template <class T>
__global__ __launch_bounds__(IX_TPB) void k_keys_bin(const T* __restrict__ col, uint32_t n,
                                                     const double* __restrict__ bounds, uint32_t nobs,
                                                     uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * IX_TPB + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = bin_locate(bounds, nobs, as_double(col[i]));
    keys[i] = j < nobs ? j : NOKEY;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (key, row) pairs, 8-bit digits.  One pass:
//   k_rs_hist     per-block digit histogram (block = RS_TILE elements)
//   scan          over the digit-major [256][nblocks] table
//   k_rs_scatter  stable scatter: each wave ranks its 64 elements with
//                 ballot match masks, waves are ordered by LDS prefix sums
constexpr int RS_TPB = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_TPB * RS_ITEMS;   // elements per block
constexpr int RS_WAVES = RS_TPB / WAVE;

__global__ __launch_bounds__(RS_TPB) void k_rs_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                    uint32_t shift, uint32_t nblocks,
                                                    uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    for (int i = threadIdx.x; i < 256; i += RS_TPB) h[i] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RS_TILE;
    for (uint32_t k = threadIdx.x; k < RS_TILE; k += RS_TPB) {
        const uint32_t i = base + k;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < 256; d += RS_TPB) hist[(uint32_t)d * nblocks + blockIdx.x] = h[d];
}

__global__ __launch_bounds__(RS_TPB) void k_rs_scatter(const uint32_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ vals, uint32_t n,
                                                       uint32_t shift, uint32_t nblocks,
                                                       const uint32_t* __restrict__ offs,
                                                       uint32_t* __restrict__ okeys,
                                                       uint32_t* __restrict__ ovals) {
    // wave w owns elements [base + w*ITEMS*64, base + (w+1)*ITEMS*64)
    __shared__ uint32_t cnt[RS_WAVES][256];
    __shared__ uint32_t boff[256];
    const int lane = (int)lane_id();
    const int w = threadIdx.x / WAVE;
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_TPB) (&cnt[0][0])[i] = 0;
    for (int d = threadIdx.x; d < 256; d += RS_TPB) boff[d] = offs[(uint32_t)d * nblocks + blockIdx.x];
    __syncthreads();
    const uint32_t wbase = blockIdx.x * RS_TILE + (uint32_t)w * RS_ITEMS * WAVE;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t dig[RS_ITEMS], rank[RS_ITEMS];
    // pass 1: ranks within the wave, counts per (wave, digit)
    #pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const uint32_t i = wbase + (uint32_t)r * WAVE + lane;
        const bool valid = i < n;
        const uint32_t d = valid ? ((keys[i] >> shift) & 255u) : 256u;
        uint64_t m = __ballot(valid);
        #pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bb = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        dig[r] = d;
        const uint32_t before = (uint32_t)__popcll(m & lt_mask);
        const uint32_t total = (uint32_t)__popcll(m);
        // the lowest lane of each digit group adds the group to the counter
        uint32_t prev = 0;
        if (valid && before == 0) { prev = cnt[w][d]; cnt[w][d] = prev + total; }
        // broadcast prev within the group: leader lane index
        const int leader = __ffsll((unsigned long long)m) - 1;
        prev = (uint32_t)__shfl((int)prev, leader < 0 ? 0 : leader, 64);
        rank[r] = prev + before;
    }
    __syncthreads();
    // exclusive prefix over waves per digit
    for (int d = threadIdx.x; d < 256; d += RS_TPB) {
        uint32_t acc = 0;
        for (int ww = 0; ww < RS_WAVES; ++ww) { const uint32_t t = cnt[ww][d]; cnt[ww][d] = acc; acc += t; }
    }
    __syncthreads();
    #pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const uint32_t i = wbase + (uint32_t)r * WAVE + lane;
        if (i < n) {
            const uint32_t d = dig[r];
            const uint32_t dst = boff[d] + cnt[w][d] + rank[r];
            okeys[dst] = keys[i];
            ovals[dst] = vals ? vals[i] : i;
        }
    }
}

// ---------------------------------------------------------------------------
// Per-bitmap encoding, chunk-parallel.  The sorted rows of bitmap b are
// [start[b], start[b+1]), ascending.  Work item c is a chunk of at most
// ENC_CR consecutive rows of one bitmap (cbm[c], cstart[c]; a bitmap without
// rows has one empty chunk), taken by one block.  Thread t takes ENC_K
// consecutive rows; the first row of each group heads a unit whose literal
// is the OR of the group's rows (at most 31, read ahead across chunk ends).
// A unit writes a 0-gap word before it when its group does not follow the
// previous unit's, and a word of its own unless it continues a run of full
// groups (the head of a run writes the run's fill, counting full groups
// ahead).  The previous unit of a thread's first unit comes from a block
// max-scan of the packed (group + 1, full) of each thread's last unit; the
// chunk's carry-in is the unit of the last row below groups*31 before the
// chunk, its fullness from the at most 31 rows of its group.  The chunk
// that ends the bitmap adds the 0-gap up to the last group.
// Pass 1 (WRITE false) counts each chunk's words, ORs the active bits
// (rows >= groups*31) into aval[b] and, for bins, each chunk's value range;
// pass 2 writes the words from wpos[c], or, for a bitmap stored
// decompressed, each unit's literal at boff[b] + group (the bitmap's words
// zeroed before), and the bitmap's last chunk its trailer.
constexpr int ENC_K = 8;
constexpr uint32_t ENC_CR = IX_TPB * ENC_K;
static_assert(ENC_CR == ENC_ROWS, "the host's chunk size");
constexpr uint64_t ENC_SKIP = ~0ull;

// Exclusive / inclusive block-wide max of 32-bit values (>= 0).
template <int NT>
__device__ __forceinline__ uint32_t block_excl_max_u32(uint32_t v, uint32_t* sh) {
    const uint32_t w = threadIdx.x / WAVE;
    uint32_t inc = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)inc, d, 64);
        if ((int)lane_id() >= d) inc = inc > o ? inc : o;
    }
    if (lane_id() == WAVE - 1) sh[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t i = 0; i < w; ++i) before = sh[i] > before ? sh[i] : before;
    uint32_t ex = (uint32_t)__shfl_up((int)inc, 1, 64);
    if (lane_id() == 0) ex = 0;
    __syncthreads();
    return ex > before ? ex : before;
}

template <int NT>
__device__ __forceinline__ uint32_t block_max_u32(uint32_t v, uint32_t* sh) {
    for (int d = 32; d >= 1; d >>= 1) { const uint32_t o = (uint32_t)__shfl_xor((int)v, d, 64); v = v > o ? v : o; }
    if (lane_id() == 0) sh[threadIdx.x / WAVE] = v;
    __syncthreads();
    uint32_t m = 0;
    for (int i = 0; i < NT / WAVE; ++i) m = sh[i] > m ? sh[i] : m;
    __syncthreads();
    return m;
}

// First index j in [s, e) with rows[j] >= r (rows ascending), else e.
__device__ __forceinline__ uint32_t lower_row(const uint32_t* rows, uint32_t s, uint32_t e, uint32_t r) {
    uint32_t lo = s, hi = e;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (rows[mid] < r) lo = mid + 1; else hi = mid;
    }
    return lo;
}

template <class T, bool WRITE>
__global__ __launch_bounds__(IX_TPB) void k_enc_chunks(EncArgs a, const T* __restrict__ col) {
    __shared__ uint32_t sh[IX_TPB / WAVE + 1];
    __shared__ uint32_t s_carry;
    __shared__ double s_mn[IX_TPB / WAVE], s_mx[IX_TPB / WAVE];
    const uint32_t lim = a.groups * 31u;
    const uint32_t* rows = a.rows;
    for (uint32_t c = blockIdx.x; c < a.nchunk; c += gridDim.x) {
        if (WRITE && a.wpos[c] == ENC_SKIP) continue;        // the same for the whole block
        const uint32_t b = a.cbm[c];
        const uint32_t s = a.start[b], e = a.start[b + 1];
        const uint32_t c0 = a.cstart[c], c1 = min(c0 + ENC_CR, e);
        const bool lit = WRITE && a.unpack != nullptr && a.unpack[b];
        if (threadIdx.x == 0) {
            // the unit of the last row below lim before the chunk
            uint32_t carry = 0;
            const uint32_t p = min(c0, lower_row(rows, s, e, lim));
            if (p > s) {
                const uint32_t g = rows[p - 1] / 31u;
                uint32_t lo = p - 1, hi = p;
                while (lo > s && rows[lo - 1] / 31u == g) --lo;
                while (hi < e && rows[hi] < lim && rows[hi] / 31u == g) ++hi;
                carry = ((g + 1u) << 1) | (hi - lo == 31u ? 1u : 0u);
            }
            s_carry = carry;
        }
        __syncthreads();
        const uint32_t carry = s_carry;
        const uint32_t i0 = c0 + threadIdx.x * (uint32_t)ENC_K;
        uint32_t ug[ENC_K];
        word_t uv[ENC_K];
        int nu = 0;
        uint32_t av = 0;
        double mn = DBL_MAX_D, mx = -DBL_MAX_D;
        uint32_t prevg = 0xFFFFFFFFu;
        if (i0 > s && i0 < e) { const uint32_t r = rows[i0 - 1]; prevg = r < lim ? r / 31u : 0xFFFFFFFEu; }
        #pragma unroll
        for (int k = 0; k < ENC_K; ++k) {
            const uint32_t i = i0 + (uint32_t)k;
            if (i >= c1) break;
            const uint32_t r = rows[i];
            if (!WRITE && col) {
                const double x = (double)col[r];
                mn = x < mn ? x : mn;
                mx = x > mx ? x : mx;
            }
            if (r >= lim) { av |= 1u << (a.anb - 1u - (r - lim)); prevg = 0xFFFFFFFEu; continue; }
            const uint32_t g = r / 31u;
            if (g == prevg) continue;
            prevg = g;
            word_t v = 0;
            for (uint32_t j = i; j < e && j < i + 31u; ++j) {
                const uint32_t r2 = rows[j];
                if (r2 / 31u != g || r2 >= lim) break;
                v |= 1u << (30u - (r2 - g * 31u));
            }
            ug[nu] = g; uv[nu] = v; ++nu;
        }
        const uint32_t mine = nu ? ((ug[nu - 1] + 1u) << 1) | (uv[nu - 1] == ALLONES ? 1u : 0u) : 0u;
        const uint32_t ex = block_excl_max_u32<IX_TPB>(mine, sh);
        const uint32_t before = ex ? ex : carry;
        const bool ends = c1 == e;                           // the chunk ends the bitmap
        if (lit) {
            // decompressed form: every unit's literal at its group
            word_t* dst = a.out + a.boff[b];
            #pragma unroll
            for (int u = 0; u < ENC_K; ++u) {
                if (u >= nu) break;
                dst[ug[u]] = uv[u];
            }
        } else {
            uint32_t cnt = 0;
            uint32_t prev = before;
            #pragma unroll
            for (int u = 0; u < ENC_K; ++u) {
                if (u >= nu) break;
                const int64_t pg = prev ? (int64_t)(prev >> 1) - 1 : -1;
                const bool pfull = prev & 1u;
                const int64_t gap = (int64_t)ug[u] - pg - 1;
                const bool merge = uv[u] == ALLONES && pfull && gap == 0;
                cnt += (gap > 0 ? 1u : 0u) + (merge ? 0u : 1u);
                prev = ((ug[u] + 1u) << 1) | (uv[u] == ALLONES ? 1u : 0u);
            }
            uint32_t tot;
            uint32_t pos = block_excl_sum<IX_TPB>(cnt, sh, &tot);
            // the 0-gap after the bitmap's last unit
            const uint32_t lastkey = ends ? block_max_u32<IX_TPB>(mine > carry ? mine : carry, sh) : 0u;
            const int64_t lg = lastkey ? (int64_t)(lastkey >> 1) - 1 : -1;
            const int64_t tgap = (int64_t)a.groups - lg - 1;
            if (!WRITE) {
                if (threadIdx.x == 0) a.ccnt[c] = tot + (ends && tgap > 0 ? 1u : 0u);
            } else {
                word_t* dst = a.out + a.wpos[c];
                prev = before;
                #pragma unroll
                for (int u = 0; u < ENC_K; ++u) {
                    if (u >= nu) break;
                    const int64_t pg = prev ? (int64_t)(prev >> 1) - 1 : -1;
                    const bool pfull = prev & 1u;
                    const int64_t gap = (int64_t)ug[u] - pg - 1;
                    const bool merge = uv[u] == ALLONES && pfull && gap == 0;
                    if (gap > 0) dst[pos++] = gap >= 2 ? (HEADER0 | (uint32_t)gap) : 0u;
                    if (!merge) {
                        if (uv[u] == ALLONES) {
                            // a run of full groups: count the following full groups
                            uint32_t L = 1;
                            uint32_t j = lower_row(rows, s, e, (ug[u] + 1u) * 31u);
                            for (;;) {
                                const uint32_t g2 = ug[u] + L;
                                if (g2 >= a.groups || j + 30u >= e) break;
                                if (rows[j] != g2 * 31u || rows[j + 30u] != g2 * 31u + 30u) break;
                                ++L; j += 31u;
                            }
                            dst[pos++] = L >= 2 ? (HEADER1 | L) : ALLONES;
                        } else {
                            dst[pos++] = uv[u];
                        }
                    }
                    prev = ((ug[u] + 1u) << 1) | (uv[u] == ALLONES ? 1u : 0u);
                }
                if (ends && threadIdx.x == 0 && tgap > 0) dst[tot] = tgap >= 2 ? (HEADER0 | (uint32_t)tgap) : 0u;
            }
        }
        if (!WRITE) {
            // active bits and the chunk's value range
            for (int d = 32; d >= 1; d >>= 1) av |= (uint32_t)__shfl_xor((int)av, d, 64);
            if (lane_id() == 0 && av) atomicOr(&a.aval[b], av);
            if (col) {
                for (int d = 32; d >= 1; d >>= 1) {
                    const double o1 = __shfl_xor(mn, d, 64), o2 = __shfl_xor(mx, d, 64);
                    mn = o1 < mn ? o1 : mn;
                    mx = o2 > mx ? o2 : mx;
                }
                if (lane_id() == 0) { s_mn[threadIdx.x / WAVE] = mn; s_mx[threadIdx.x / WAVE] = mx; }
                __syncthreads();
                if (threadIdx.x == 0) {
                    double lo = DBL_MAX_D, hi = -DBL_MAX_D;
                    for (int w = 0; w < IX_TPB / WAVE; ++w) { lo = s_mn[w] < lo ? s_mn[w] : lo; hi = s_mx[w] > hi ? s_mx[w] : hi; }
                    a.cmin[c] = lo; a.cmax[c] = hi;
                }
            }
        } else if (ends && threadIdx.x == 0) {
            // the trailer: active value and bit count (or the 0 of no active bits)
            word_t* t = a.out + a.boff[b] + a.bnw[b];
            if (a.anb) { t[0] = a.aval[b]; t[1] = a.anb; }
            else t[0] = 0u;
        }
        __syncthreads();
    }
}

// Tile offset table, chunk-parallel.  Chunk c is code words
// [tw0[c], min(tw0[c] + TOFF_CH, nw[b])) of bitmap b = tb[c]; word indices
// in the table are absolute (into the serialized array).  Pass 1 sums each
// chunk's word lengths; with the chunks' start groups (a per-bitmap scan on
// the host) pass 2 gives every word its span, and the word whose span holds
// group t*tg writes row t of its bitmap.  The tail pass writes the rows of
// tiles past a bitmap's last group ({end, groups}) and the closing row
// ntiles ({end, 1 when the bitmap has words and a trailer}).
constexpr uint32_t TOFF_PER = 16;
constexpr uint32_t TOFF_CH = IX_TPB * TOFF_PER;
static_assert(TOFF_CH == TOFF_WORDS, "the host's chunk size");

__global__ __launch_bounds__(IX_TPB) void k_toff_len(const uint32_t* __restrict__ words,
                                                     const uint64_t* __restrict__ off,
                                                     const uint32_t* __restrict__ nw,
                                                     const uint32_t* __restrict__ tb,
                                                     const uint32_t* __restrict__ tw0, uint32_t nch,
                                                     uint32_t* __restrict__ clen) {
    __shared__ uint32_t sh[IX_TPB / WAVE + 1];
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint32_t b = tb[c];
        const uint32_t* w = words + off[b];
        const uint32_t w0 = tw0[c], w1 = min(w0 + TOFF_CH, nw[b]);
        uint32_t l = 0;
        for (uint32_t i = w0 + threadIdx.x; i < w1; i += IX_TPB) l += word_len(w[i]);
        uint32_t tot;
        (void)block_excl_sum<IX_TPB>(l, sh, &tot);
        if (threadIdx.x == 0) clen[c] = tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(IX_TPB) void k_toff_rows(const uint32_t* __restrict__ words,
                                                      const uint64_t* __restrict__ off,
                                                      const uint32_t* __restrict__ nw,
                                                      const uint32_t* __restrict__ tb,
                                                      const uint32_t* __restrict__ tw0,
                                                      const uint32_t* __restrict__ cg0, uint32_t nch,
                                                      uint32_t nbm, uint32_t ntiles, uint32_t tg,
                                                      uint2* __restrict__ toff) {
    __shared__ uint32_t sh[IX_TPB / WAVE + 1];
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint32_t b = tb[c];
        const uint32_t base = (uint32_t)off[b];
        const uint32_t w0 = tw0[c], w1 = min(w0 + TOFF_CH, nw[b]);
        const uint32_t i0 = w0 + threadIdx.x * TOFF_PER;
        uint32_t len[TOFF_PER];
        uint32_t sum = 0;
        #pragma unroll
        for (uint32_t k = 0; k < TOFF_PER; ++k) {
            len[k] = i0 + k < w1 ? word_len(words[base + i0 + k]) : 0u;
            sum += len[k];
        }
        uint32_t tot;
        uint32_t pos = cg0[c] + block_excl_sum<IX_TPB>(sum, sh, &tot);
        #pragma unroll
        for (uint32_t k = 0; k < TOFF_PER; ++k) {
            if (len[k]) {
                const uint32_t end = pos + len[k];
                for (uint32_t t = (pos + tg - 1) / tg; t < ntiles && (uint64_t)t * tg < end; ++t)
                    toff[(size_t)t * nbm + b] = make_uint2(base + i0 + k, pos);
            }
            pos += len[k];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(IX_TPB) void k_toff_tail(const uint64_t* __restrict__ off,
                                                      const uint32_t* __restrict__ nw,
                                                      const uint32_t* __restrict__ total, uint32_t nbm,
                                                      uint32_t ntiles, uint32_t tg, uint32_t groups,
                                                      uint2* __restrict__ toff) {
    const uint32_t b = blockIdx.x * IX_TPB + threadIdx.x;
    if (b >= nbm) return;
    const uint32_t end = (uint32_t)off[b] + nw[b];
    toff[(size_t)ntiles * nbm + b] = make_uint2(end, off[b + 1] == off[b] ? 0u : 1u);
    for (uint32_t t = (total[b] + tg - 1) / tg; t < ntiles; ++t)
        toff[(size_t)t * nbm + b] = make_uint2(end, groups);
}

// ---------------------------------------------------------------------------
// Range evaluation, LDS plan.  One block per tile of QT groups; the tile
// result R[g] (plus the active slot at g == groups, left-aligned) is the OR
// of bitmaps in up to two ranges, optionally complemented, then OR-ed with
// the rows of the candidate (edge) bitmaps whose raw values satisfy the
// range.
template <class T>
__device__ __forceinline__ bool in_range(const QueryArgs& q, T x) {
    if (q.is_int) {
        const int64_t v = (int64_t)x;
        bool ok = true;
        if (q.lop == OPC_LT) ok = v > q.ilo; else if (q.lop == OPC_LE) ok = v >= q.ilo;
        if (q.rop == OPC_LT) ok = ok && v < q.ihi; else if (q.rop == OPC_LE) ok = ok && v <= q.ihi;
        return ok;
    }
    const double v = (double)x;
    bool ok = true;
    if (q.lop == OPC_LT) ok = q.lo < v; else if (q.lop == OPC_LE) ok = q.lo <= v;
    if (q.rop == OPC_LT) ok = ok && v < q.hi; else if (q.rop == OPC_LE) ok = ok && v <= q.hi;
    return ok;
}

// ---------------------------------------------------------------------------
// Tiled copy of the bitmaps for the literal plan.  For every tile t of QT
// groups and bitmap b, segment (t, b) holds the code words of b that overlap
// the tile with fills clipped to it, so that every segment of a bitmap with
// words covers exactly the tile's groups (plus, in the last tile, the active
// word appended left-aligned as one more literal group).  Segments are laid
// out tile-major: tw[TW[t*nobs + b] .. TW[t*nobs + b + 1]).

// Code words of bitmap b overlapping tile t: [s, e) with s's start group p
// (from the tile offset table).
__device__ __forceinline__ void tile_segment(const uint2* toff, uint32_t nobs, uint32_t ntiles,
                                             uint32_t groups, uint32_t t, uint32_t b, uint32_t& s,
                                             uint32_t& e, uint32_t& p) {
    const uint2 F0 = toff[(size_t)t * nobs + b];
    const uint2 F1 = toff[(size_t)(t + 1) * nobs + b];
    const uint32_t g1 = min((t + 1) * (uint32_t)QT, groups);
    s = F0.x; p = F0.y;
    e = (t + 1 == ntiles) ? F1.x : F1.x + (F1.y < g1 ? 1u : 0u);
}

__global__ __launch_bounds__(IX_TPB) void k_tseg_len(const uint2* __restrict__ toff, uint32_t nobs,
                                                     uint32_t ntiles, uint32_t groups, uint32_t anb,
                                                     uint32_t* __restrict__ len) {
    const uint32_t idx = blockIdx.x * IX_TPB + threadIdx.x;
    if (idx > ntiles * nobs) return;
    if (idx == ntiles * nobs) { len[idx] = 0u; return; }     // total after the scan
    const uint32_t t = idx / nobs, b = idx - t * nobs;
    uint32_t s, e, p;
    tile_segment(toff, nobs, ntiles, groups, t, b, s, e, p);
    uint32_t L = e - s;
    if (t + 1 == ntiles && anb > 0 && toff[(size_t)ntiles * nobs + b].y) ++L;   // active word
    len[idx] = L;
}

// One wave per segment: copy with the first and last fill clipped.
__global__ __launch_bounds__(IX_TPB) void k_tseg_fill(const uint32_t* __restrict__ words,
                                                      const uint2* __restrict__ toff,
                                                      const uint32_t* __restrict__ TW, uint32_t nobs,
                                                      uint32_t ntiles, uint32_t groups, uint32_t anb,
                                                      uint32_t* __restrict__ tw) {
    // segments in grid-stride order: a grid of one wave per segment can
    // exceed 2^32 work-items (the dispatch's 32-bit grid size)
    const uint32_t nseg = ntiles * nobs;
    for (uint32_t idx = blockIdx.x * (IX_TPB / WAVE) + threadIdx.x / WAVE; idx < nseg;
         idx += gridDim.x * (IX_TPB / WAVE)) {
    const int lane = (int)lane_id();
    const uint32_t t = idx / nobs, b = idx - t * nobs;
    uint32_t s, e, p;
    tile_segment(toff, nobs, ntiles, groups, t, b, s, e, p);
    const uint32_t g0 = t * (uint32_t)QT, g1 = min(g0 + (uint32_t)QT, groups);
    const uint32_t base = TW[idx];
    const uint32_t L = e - s;
    uint32_t carry = p;
    for (uint32_t o = 0; o < L; o += WAVE) {
        const uint32_t i = o + (uint32_t)lane;
        const word_t x = i < L ? words[s + i] : 0u;
        const uint32_t len = i < L ? word_len(x) : 0u;
        const uint32_t inc = wave_incl_sum(len);
        const uint32_t pos = carry + inc - len;
        if (i < L) {
            word_t y = x;
            if (x >= HEADER0) {
                const uint32_t lo = pos > g0 ? pos : g0;
                const uint32_t hi = pos + len < g1 ? pos + len : g1;
                y = (x & HEADER1) | (hi - lo);
            }
            tw[base + i] = y;
        }
        carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, WAVE - 1);
    }
    if (lane == 0 && t + 1 == ntiles && anb > 0) {
        const uint2 E = toff[(size_t)ntiles * nobs + b];
        if (E.y) tw[base + L] = words[E.x] << (31u - anb);
    }
    }
}


// Literal range plan.  One block of 192 or 256 threads owns tile t: ng = QT
// groups (in the last tile the remaining groups plus the active slot).  The
// segments of a bitmap range are contiguous in the tiled copy and each
// covers exactly ng groups, so a word's group within the tile is the
// running sum of the word lengths from the start of any segment, modulo
// ng.  The OR set of the tile is one or two contiguous word streams; the
// block splits them into per-wave pieces of equal word counts at segment
// boundaries (an LDS table of chunk starts), and each wave streams its
// piece 64 words a sub-round with a DPP prefix sum of the word lengths,
// QB_U sub-rounds per round and the next round's loads issued before the
// current round is applied.  Literal words are OR-ed into the LDS tile R;
// 1-fills add +1/-1 at their ends to the LDS coverage array cov, whose
// prefix sums mark all-one groups.  Edge bitmaps stream one sub-round at a
// time with every row checked against the raw column.
constexpr int QB_U = 2;                      // sub-rounds (16-byte loads) per round
constexpr int CJ_U = 2;                      // the same inside the conjunction kernel
constexpr int CT_MAX = 128;                  // chunks of the split table
constexpr int CT_WORDS = CT_MAX + 8;         // LDS words of the split table

template <class T>
__device__ __forceinline__ word_t check_group(const QueryArgs& q, const T* col, uint32_t g, word_t bits) {
    word_t m = 0;
    while (bits) {
        const uint32_t hb = 31u - __clz(bits);
        const uint32_t row = g * 31u + (30u - hb);
        if (in_range(q, col[row])) m |= 1u << hb;
        bits &= ~(1u << hb);
    }
    return m;
}

// Applies one sub-round of 64 words (lane's word w, act = lane holds a
// word) at positions carry + prefix (mod ng); returns the new carry.
template <class T, bool CHECK, bool POW2>
__device__ __forceinline__ uint32_t apply_words(const QueryArgs& q, word_t w, bool act, uint32_t carry,
                                                uint32_t g0, uint32_t ng, word_t* R, uint32_t* cov, const T* col) {
    const uint32_t len = act ? word_len(w) : 0u;
    const uint32_t inc = wave_incl_sum(len);
    uint32_t rel = carry + inc - len;
    rel = POW2 ? (rel & (uint32_t)(QT - 1)) : (rel % ng);
    const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)inc, WAVE - 1);
    carry = POW2 ? ((carry + tot) & (uint32_t)(QT - 1)) : ((carry + tot) % ng);
    if (act) {
        if (w < HEADER0) {
            if (w) {
                const word_t m = CHECK ? check_group<T>(q, col, g0 + rel, w) : w;
                if (m) atomicOr(&R[rel], m);
            }
        } else if (w >= HEADER1) {
            if (!CHECK) {
                atomicAdd(&cov[rel], 1u);
                atomicSub(&cov[rel + len], 1u);
            } else {
                for (uint32_t k = 0; k < len; ++k) {
                    const word_t m = check_group<T>(q, col, g0 + rel + k, ALLONES);
                    if (m) atomicOr(&R[rel + k], m);
                }
            }
        }
    }
    return carry;
}

// Edge bitmap words [ws, we) (one segment), every row checked, by one wave.
template <class T, bool POW2>
__device__ void stream_checked(const QueryArgs& q, uint32_t ws, uint32_t we, uint32_t g0, uint32_t ng,
                               word_t* R, uint32_t* cov, const T* col) {
    const int lane = (int)lane_id();
    uint32_t carry = 0;
    for (uint32_t base = ws; base < we; base += WAVE) {
        const uint32_t i = base + (uint32_t)lane;
        const word_t w = i < we ? q.tw[i] : 0u;
        carry = apply_words<T, true, POW2>(q, w, i < we, carry, g0, ng, R, cov, col);
    }
}

// Applies word w (valid: in the stream) at tile position p.
__device__ __forceinline__ void apply_one(word_t w, bool valid, uint32_t p, uint32_t len, word_t* R,
                                          uint32_t* cov) {
    if (valid && w != 0u && w < HEADER0) atomicOr(&R[p], w);
    else if (valid && w >= HEADER1) {
        atomicAdd(&cov[p], 1u);
        atomicSub(&cov[p + len], 1u);
    }
}

// Loads of one round: every load is issued (lanes past the stream read its
// last group again, and apply_group4 drops words outside [ws, we)), so the
// loads are straight-line code and each wait counts only the loads issued
// after the ones it needs.
template <int U>
__device__ __forceinline__ void load_round4(const uint4* tw4, uint32_t base4, uint32_t end4, uint4 (&x)[U]) {
    const uint32_t lane = lane_id();
    #pragma unroll
    for (int u = 0; u < U; ++u) x[u] = tw4[min(base4 + (uint32_t)u * WAVE + lane, end4 - 1u)];
}

// One sub-round: lane l applies the four words of 16-byte group g4 + l at
// carry + their prefix sum over the wave (mod ng); returns the new carry.
// FULL: every word of the sub-round lies in [ws, we) (no checks).
template <bool POW2, bool FULL, bool TRACK>
__device__ __forceinline__ uint32_t apply_group4(const uint4& x, uint32_t g4, uint32_t ws, uint32_t we,
                                                 uint32_t carry, uint32_t ng, word_t* R, uint32_t* cov,
                                                 uint32_t& fl) {
    const uint32_t j = (g4 + lane_id()) << 2;          // word index of x.x
    const bool v0 = FULL || (j >= ws && j < we), v1 = FULL || (j + 1 >= ws && j + 1 < we);
    const bool v2 = FULL || (j + 2 >= ws && j + 2 < we), v3 = FULL || (j + 3 >= ws && j + 3 < we);
    const uint32_t l0 = v0 ? word_len(x.x) : 0u, l1 = v1 ? word_len(x.y) : 0u;
    const uint32_t l2 = v2 ? word_len(x.z) : 0u, l3 = v3 ? word_len(x.w) : 0u;
    const uint32_t sum = l0 + l1 + l2 + l3;
    const uint32_t inc = wave_incl_sum(sum);
    uint32_t p0 = carry + inc - sum, p1 = p0 + l0, p2 = p1 + l1, p3 = p2 + l2;
    if (POW2) {
        p0 &= (uint32_t)(QT - 1); p1 &= (uint32_t)(QT - 1);
        p2 &= (uint32_t)(QT - 1); p3 &= (uint32_t)(QT - 1);
    } else {
        p0 %= ng; p1 %= ng; p2 %= ng; p3 %= ng;
    }
    const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)inc, WAVE - 1);
    apply_one(x.x, v0, p0, l0, R, cov);
    apply_one(x.y, v1, p1, l1, R, cov);
    apply_one(x.z, v2, p2, l2, R, cov);
    apply_one(x.w, v3, p3, l3, R, cov);
    if (TRACK)
        fl |= (v0 && x.x >= HEADER1) | (v1 && x.y >= HEADER1) | (v2 && x.z >= HEADER1) | (v3 && x.w >= HEADER1);
    return POW2 ? ((carry + tot) & (uint32_t)(QT - 1)) : ((carry + tot) % ng);
}

// Words [ws, we) (whole segments) by one wave.  Lane l of sub-round u takes
// the four words of 16-byte group base4 + 64u + l (words outside [ws, we)
// count as absent), so one prefix sum over the wave places 256 words; a
// round of U sub-rounds of loads is in flight while the previous round is
// applied.  The tiled copy is padded so that whole groups can be read.
template <bool POW2, int U, bool TRACK>
__device__ __forceinline__ uint32_t apply_round4(const uint4 (&x)[U], uint32_t b4, uint32_t end4, uint32_t full0,
                                                 uint32_t full1, uint32_t ws, uint32_t we, uint32_t carry,
                                                 uint32_t ng, word_t* R, uint32_t* cov, uint32_t& fl) {
    #pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t g4 = b4 + (uint32_t)u * WAVE;
        if (g4 >= end4) break;
        if (g4 >= full0 && g4 + WAVE <= full1)
            carry = apply_group4<POW2, true, TRACK>(x[u], g4, ws, we, carry, ng, R, cov, fl);
        else
            carry = apply_group4<POW2, false, TRACK>(x[u], g4, ws, we, carry, ng, R, cov, fl);
    }
    return carry;
}

template <bool POW2, int U, bool TRACK>
__device__ __forceinline__ void stream_words(const QueryArgs& q, uint32_t ws, uint32_t we, uint32_t ng,
                                             word_t* R, uint32_t* cov, uint32_t& fl) {
    if (ws >= we) return;
    const uint4* tw4 = reinterpret_cast<const uint4*>(q.tw);
    const uint32_t a4 = ws >> 2, end4 = (we + 3u) >> 2;
    const uint32_t full0 = (ws + 3u) >> 2, full1 = we >> 2;   // groups wholly inside
    uint32_t carry = 0;
    // two batches of U rounds alternate: one in flight while the other is
    // applied
    uint4 xa[U], xb[U];
    load_round4<U>(tw4, a4, end4, xa);
    for (uint32_t b4 = a4; b4 < end4; b4 += 2u * U * WAVE) {
        load_round4<U>(tw4, b4 + U * WAVE, end4, xb);
        carry = apply_round4<POW2, U, TRACK>(xa, b4, end4, full0, full1, ws, we, carry, ng, R, cov, fl);
        load_round4<U>(tw4, b4 + 2u * U * WAVE, end4, xa);
        carry = apply_round4<POW2, U, TRACK>(xb, b4 + U * WAVE, end4, full0, full1, ws, we, carry, ng, R, cov, fl);
    }
}

// A single bitmap whose segment holds one word per group (literal words
// and count-1 fills: a dense bitmap, such as a coarse bitmap of a fuzz
// index) is copied by the whole block, each thread owning groups
// k = tid (mod NT), with no atomics and no coverage pass; the streaming
// path would give it to one wave.

// The OR set of the tile: stream 1 = the segments of bitmaps [r0, r1),
// stream 2 = those of [r2, r3).  The block reads the starts of up to CT_MAX
// chunks of consecutive bitmaps into ctab; wave w streams the chunks whose
// start lies in the w-th quarter of the words.  With DENSE, R[0, QT] is
// cleared here while the chunk starts are in flight (else by the caller).
// Called by every thread of the block; returns true when the dense form
// above did the work, and sets fl in a thread that recorded a 1-fill in
// cov.
template <class T, bool POW2, int NT, int U, bool DENSE>
__device__ bool tile_ranges(const QueryArgs& q, uint32_t tile, uint32_t g0, uint32_t ng, word_t* R,
                            uint32_t* cov, const T* col, uint32_t* ctab, uint32_t& fl) {
    constexpr int NW = NT / WAVE;
    static_assert(2 * NT >= CT_MAX + 2, "two chunk starts per thread");
    const uint32_t len1 = q.r1 - q.r0, len2 = q.r3 - q.r2;
    if (!DENSE && len1 + len2 == 0) return true;
    const uint32_t cb = max(16u, (len1 + len2 + CT_MAX - 1) / CT_MAX);   // bitmaps per chunk
    const uint32_t n1 = (len1 + cb - 1) / cb, n2 = (len2 + cb - 1) / cb;
    const uint32_t nc = n1 + n2;                                        // <= CT_MAX + 2
    const uint32_t* TWt = q.TW + (size_t)tile * q.nobs;
    uint32_t* st = ctab + CT_MAX + 2;        // s1, e1, s2, e2
    if (DENSE) {
        // R cleared while the chunk starts are in flight
        const uint32_t ka0 = threadIdx.x, ka1 = threadIdx.x + NT;
        uint32_t t0 = 0, t1 = 0;
        if (ka0 < nc) t0 = TWt[ka0 < n1 ? q.r0 + ka0 * cb : q.r2 + (ka0 - n1) * cb];
        if (ka1 < nc) t1 = TWt[ka1 < n1 ? q.r0 + ka1 * cb : q.r2 + (ka1 - n1) * cb];
        if (threadIdx.x == NT - 1) {         // the last wave, whose chunk loads are the fewest
            st[0] = len1 ? TWt[q.r0] : 0u;
            st[1] = len1 ? TWt[q.r1] : 0u;
            st[2] = len2 ? TWt[q.r2] : 0u;
            st[3] = len2 ? TWt[q.r3] : 0u;
        }
        for (uint32_t k = threadIdx.x; k <= (uint32_t)QT; k += NT) R[k] = 0u;
        if (ka0 < nc) ctab[ka0] = t0;
        if (ka1 < nc) ctab[ka1] = t1;
    } else {
        // the caller cleared R
        for (uint32_t k = threadIdx.x; k < nc; k += NT)
            ctab[k] = TWt[k < n1 ? q.r0 + k * cb : q.r2 + (k - n1) * cb];
        if (threadIdx.x == 0) {
            st[0] = len1 ? TWt[q.r0] : 0u;
            st[1] = len1 ? TWt[q.r1] : 0u;
            st[2] = len2 ? TWt[q.r2] : 0u;
            st[3] = len2 ? TWt[q.r3] : 0u;
        }
    }
    __syncthreads();
    if (len1 + len2 == 0) return true;
    const uint32_t s1 = st[0], e1 = st[1], s2 = st[2], e2 = st[3];
    const uint32_t W1 = e1 - s1, W = W1 + (e2 - s2);
    if (DENSE && len1 + len2 == 1 && W == ng) {
        const uint32_t s = len1 ? s1 : s2;
        for (uint32_t i = threadIdx.x; i < ng; i += NT) {
            const word_t w = q.tw[s + i];
            R[i] = w < HEADER0 ? w : (w >= HEADER1 ? ALLONES : 0u);
        }
        return true;
    }
    // word offset of chunk k in the concatenated streams
    auto cstart = [&](uint32_t k) -> uint32_t {
        if (k >= nc) return W;
        return k < n1 ? ctab[k] - s1 : W1 + (ctab[k] - s2);
    };
    // first chunk whose start is >= target (binary search over the table)
    auto first_ge = [&](uint32_t target) -> uint32_t {
        uint32_t lo = 0, hi = nc;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cstart(mid) < target) lo = mid + 1; else hi = mid;
        }
        return lo;
    };
    const uint32_t w = threadIdx.x / WAVE;
    const uint32_t ka = w == 0 ? 0u : first_ge((uint32_t)(((uint64_t)W * w) / NW));
    const uint32_t kb = w + 1 == NW ? nc : first_ge((uint32_t)(((uint64_t)W * (w + 1)) / NW));
    if (ka >= kb) return false;
    // the piece in stream 1, then the piece in stream 2
    if (ka < n1) {
        const uint32_t we = kb >= n1 ? e1 : ctab[kb];
        stream_words<POW2, U, DENSE>(q, ctab[ka], we, ng, R, cov, fl);
    }
    if (kb > n1) {
        const uint32_t ws = ka > n1 ? ctab[ka] : s2;
        const uint32_t we = kb >= nc ? e2 : ctab[kb];
        stream_words<POW2, U, DENSE>(q, ws, we, ng, R, cov, fl);
    }
    return false;
}

// Literal plan of one term over tile t into R[0, ng), by a block of NT
// threads.  cov[0, QT] is zero on entry.  With DENSE, R is cleared here
// and the coverage pass clears what the term wrote to cov; without, the
// caller clears both before each call.  The call ends with a barrier.  red needs NT/64 + 1 words.  DENSE enables the block-wide copy
// of a dense single-bitmap term (off in the range-batch kernel, where its
// registers would cost a wave per SIMD).
template <class T, int NT, int U, bool DENSE = true, bool EDGES = true>
__device__ __forceinline__ void lit_tile(const QueryArgs& q, uint32_t tile, uint32_t g0, uint32_t ng, bool last, word_t* R,
                         uint32_t* cov, uint32_t* red, uint32_t* ctab) {
    const T* col = (const T*)q.col;
    const int w = threadIdx.x / WAVE;
    uint32_t fl = 0;
    if (ng) {
        if (last) tile_ranges<T, false, NT, U, DENSE>(q, tile, g0, ng, R, cov, col, ctab, fl);
        else tile_ranges<T, true, NT, U, DENSE>(q, tile, g0, ng, R, cov, col, ctab, fl);
    }
    // coverage prefix sums: groups under some 1-fill are all ones; the pass
    // leaves cov zero.  With DENSE it runs only when some 1-fill was
    // recorded (a block-wide vote at the barrier).
    bool any_fill = true;
    if (DENSE) any_fill = __syncthreads_or((int)fl) != 0;
    else __syncthreads();
    if (any_fill) {
        constexpr int PER = (QT + NT - 1) / NT;
        const uint32_t k0 = threadIdx.x * PER;
        uint32_t c[PER], sum = 0;
        #pragma unroll
        for (int i = 0; i < PER; ++i) {
            c[i] = k0 + i < (uint32_t)QT ? cov[k0 + i] : 0u;
            if (DENSE && k0 + i < (uint32_t)QT) cov[k0 + i] = 0u;
            sum += c[i];
        }
        if (DENSE && threadIdx.x == 0) cov[QT] = 0u;   // the end mark of a fill reaching the tile's end
        uint32_t tot;
        uint32_t run = block_excl_sum<NT>(sum, red, &tot);
        #pragma unroll
        for (int i = 0; i < PER; ++i) {
            run += c[i];
            if (run != 0u && k0 + i < (uint32_t)QT) R[k0 + i] = ALLONES;
        }
    }
    __syncthreads();
    if (q.flip) {
        for (uint32_t k = threadIdx.x; k < ng; k += NT) {
            const uint32_t g = g0 + k;
            const word_t mask = g < q.groups ? ALLONES : (((1u << q.anb) - 1u) << (31u - q.anb));
            R[k] ^= mask;
        }
        __syncthreads();
    }
    if (EDGES && w < 2 && ng) {
        const uint32_t cb = w == 0 ? q.cand0 : q.cand1;
        if (cb != NOKEY) {
            const uint32_t* TWt = q.TW + (size_t)tile * q.nobs;
            if (last) stream_checked<T, false>(q, TWt[cb], TWt[cb + 1], g0, ng, R, cov, col);
            else stream_checked<T, true>(q, TWt[cb], TWt[cb + 1], g0, ng, R, cov, col);
        }
    }
    __syncthreads();
}

// Scan plan of one term over tile t into R[0, ng): every wave takes 64
// groups at a time (31 coalesced loads of 64 values, ballots staged in
// LDS, each lane assembles its group).  bal needs NT/64 x 32 words of 64
// bits.  Ends with a barrier.
// Ballots of the 1984 rows rbase .. rbase+1983 (64 groups) into bal[0..30]
// by one wave, B loads in flight at a time (all 31 for the stand-alone
// scan; fewer inside the conjunction kernel, whose registers are shared
// with the other plans).
template <class T, int B>
__device__ __forceinline__ void scan_rows64(const QueryArgs& q, const T* col, uint32_t rbase, uint32_t nrows,
                                            uint64_t* bal) {
    const int lane = (int)lane_id();
    #pragma unroll
    for (int k0 = 0; k0 < 31; k0 += B) {
        T v[B];
        #pragma unroll
        for (int k = 0; k < B; ++k) {
            const uint32_t row = rbase + (uint32_t)(k0 + k) * 64u + lane;
            v[k] = (k0 + k < 31 && row < nrows) ? col[row] : T(0);
        }
        #pragma unroll
        for (int k = 0; k < B; ++k) use_loaded(v[k]);
        #pragma unroll
        for (int k = 0; k < B; ++k) {
            if (k0 + k >= 31) break;
            const uint32_t row = rbase + (uint32_t)(k0 + k) * 64u + lane;
            const uint64_t m = __ballot(row < nrows && in_range(q, v[k]));
            if (lane == 0) bal[k0 + k] = m;
        }
    }
}

template <class T, int NT, int B = 8>
__device__ __forceinline__ void scan_tile(const QueryArgs& q, uint32_t g0, uint32_t ng, word_t* R, uint64_t (*bal)[32]) {
    const T* col = (const T*)q.col;
    const int lane = (int)lane_id();
    const int w = threadIdx.x / WAVE;
    const uint32_t nrows = q.groups * 31u + q.anb;
    for (uint32_t k0 = (uint32_t)w * 64u; k0 < ng; k0 += (NT / WAVE) * 64u) {
        const uint32_t rbase = (g0 + k0) * 31u;
        scan_rows64<T, B>(q, col, rbase, nrows, bal[w]);
        __builtin_amdgcn_wave_barrier();
        const uint32_t k = k0 + (uint32_t)lane;
        if (k < ng) {
            const uint32_t bit0 = (uint32_t)lane * 31u;
            const uint32_t kk = bit0 >> 6, s0 = bit0 & 63u;
            uint64_t lo = bal[w][kk] >> s0;
            if (s0 > 33u && kk + 1 < 31u) lo |= bal[w][kk + 1] << (64u - s0);
            word_t v = __brev((uint32_t)(lo & 0x7FFFFFFFull)) >> 1;
            if (g0 + k == q.groups) v &= (((1u << q.anb) - 1u) << (31u - q.anb));
            R[k] = v;
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
}

template <class T, int NT>
__global__ __launch_bounds__(NT) void k_query_lit(QueryArgs q) {
    constexpr int NWV = NT / WAVE;
    __shared__ word_t R[QT + 1];
    __shared__ uint32_t cov[QT + 1];
    __shared__ uint32_t red[NWV + 1];
    __shared__ uint32_t ctab[CT_WORDS];
    const uint32_t tile = blockIdx.x;
    const bool last = tile + 1 == q.ntiles;
    const uint32_t g0 = tile * (uint32_t)QT;
    const uint32_t ng = last ? (q.groups - g0) + (q.anb ? 1u : 0u) : (uint32_t)QT;
    const int w = threadIdx.x / WAVE;
    for (uint32_t k = threadIdx.x; k <= QT; k += NT) { R[k] = 0u; cov[k] = 0u; }
    __syncthreads();
    lit_tile<T, NT, QB_U>(q, tile, g0, ng, last, R, cov, red, ctab);
    uint32_t ones = 0;
    for (uint32_t k = threadIdx.x; k < ng; k += NT) {
        const word_t v = R[k];
        q.out[g0 + k] = v;
        ones += __popc(v);
    }
    ones = wave_sum(ones);
    if (lane_id() == 0) red[w] = ones;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < NWV; ++i) t += red[i];
        q.part[tile] = t;
    }
}

// Compressed plan of one range (sumBins leaving its result compressed, the
// priority-queue branch of src/index.cpp:6687-6760): block t merges the OR
// set's segments of tile t in LDS as the literal plan does, then writes the
// tile's words of the canonical WAH result directly.  Every whole group is a
// token of its class (all 0, all 1, mixed); thread i summarises its QT/NT
// consecutive groups, the summaries are scanned over the block and carried
// from tile to tile by the look-back (tile t is block t), and every word is
// written at its slot: a literal by its group, a run by the group that
// closes it, at its head's slot (a run may start tiles before).  No literal
// result reaches HBM.  The active slot (the last tile's last group when
// anb > 0) is not a code word: its value goes to o.ctl[2].
struct IxWriter {
    word_t* out;
    uint32_t* outck;
    __device__ __forceinline__ void operator()(uint32_t slot, word_t w, uint32_t head) {
        out[slot] = w;
        if ((slot & 31u) == 0) outck[slot >> 5] = head;
    }
};

template <class T, int NT>
__global__ __launch_bounds__(NT) void k_query_wah(QueryArgs q, WahOut o) {
    constexpr int NWV = NT / WAVE;
    constexpr int PER = QT / NT;
    __shared__ word_t R[QT + 1];
    __shared__ uint32_t cov[QT + 1];
    __shared__ uint32_t red[NWV + 1];
    __shared__ uint32_t ctab[CT_WORDS];
    __shared__ Summary sbuf[NWV];
    __shared__ Summary s_pre;
    const uint32_t tile = blockIdx.x;
    const bool last = tile + 1 == q.ntiles;
    const uint32_t g0 = tile * (uint32_t)QT;
    const uint32_t ng = last ? (q.groups - g0) + (q.anb ? 1u : 0u) : (uint32_t)QT;
    const uint32_t nw = last ? q.groups - g0 : (uint32_t)QT;          // whole groups
    const int w = threadIdx.x / WAVE;
    for (uint32_t k = threadIdx.x; k <= QT; k += NT) { R[k] = 0u; cov[k] = 0u; }
    __syncthreads();
    lit_tile<T, NT, QB_U>(q, tile, g0, ng, last, R, cov, red, ctab);
    const uint32_t k0 = threadIdx.x * (uint32_t)PER;
    word_t v[PER];
    uint32_t ones = 0;
    CountSink cs;
    cs.init();
    #pragma unroll
    for (int i = 0; i < PER; ++i) {
        v[i] = k0 + i < ng ? R[k0 + i] : 0u;
        ones += __popc(v[i]);
        if (k0 + i < nw) emit_lit(cs, v[i]);
    }
    ones = wave_sum(ones);
    if (lane_id() == 0) red[w] = ones;
    Summary agg;
    const Summary texcl = block_excl_summary2<NT>(cs.s, sbuf, &agg);
    if (threadIdx.x < WAVE) {
        const Summary pr = prefix_summary_seg(o.status, o.status + q.ntiles, tile, o.tag, agg);
        if (threadIdx.x == 0) {
            s_pre = pr;
            uint32_t t = 0;
            for (int i = 0; i < NWV; ++i) t += red[i];
            q.part[tile] = t;
        }
    }
    __syncthreads();
    const Summary pre = combine(s_pre, texcl);
    IxWriter iw{o.out, o.outck};
    WriteSink<IxWriter> ws;
    ws.init(&iw, pre, g0 + k0);
    #pragma unroll
    for (int i = 0; i < PER; ++i)
        if (k0 + i < nw) emit_lit(ws, v[i]);
    // the thread holding the last whole group (thread 0 when the last tile
    // has none) closes the last run and records the word count
    if (last && ((nw == 0 && threadIdx.x == 0) || (k0 < nw && nw <= k0 + PER))) {
        ws.finish();
        const uint32_t total = combine(pre, cs.s).count;
        o.ctl[1] = total;
        o.outck[(total + 31u) >> 5] = q.groups;
        o.ctl[2] = q.anb ? R[nw] : 0u;
    }
}

// ---------------------------------------------------------------------------
// Conjunctions (countQuery::doEvaluate, LOGICAL_AND, src/countQuery.cpp:
// 953-963): the result of a query is the AND of its terms' results.  One
// persistent kernel takes work items (query, TPI consecutive tiles) in
// grid-stride order; per tile the block evaluates the first term into R,
// every further term into R2 and R &= R2, and stops early once the tile is
// empty.  Each term runs the plan chosen for it on the host (literal
// through the tiled copy of its index, or the raw scan of its column).
template <int NT, bool LIT = false>
__device__ __forceinline__ void term_tile(const QueryArgs& q, uint32_t tile, uint32_t g0, uint32_t ng, bool last, word_t* R,
                          uint32_t* cov, uint32_t* red, uint64_t (*bal)[32], uint32_t* ctab) {
    if (LIT) {
        // literal plan, no edge bitmaps: the column type does not matter
        lit_tile<uint32_t, NT, CJ_U, true, false>(q, tile, g0, ng, last, R, cov, red, ctab);
        return;
    }
    if (q.plan == PLAN_SCAN_D) {
        switch (q.type) {
        case COL_I32: scan_tile<int32_t, NT>(q, g0, ng, R, bal); break;
        case COL_U32: scan_tile<uint32_t, NT>(q, g0, ng, R, bal); break;
        case COL_U16: scan_tile<uint16_t, NT>(q, g0, ng, R, bal); break;
        case COL_F32: scan_tile<float, NT>(q, g0, ng, R, bal); break;
        default: scan_tile<double, NT>(q, g0, ng, R, bal); break;
        }
        return;
    }
    switch (q.type) {
    case COL_I32: lit_tile<int32_t, NT, CJ_U>(q, tile, g0, ng, last, R, cov, red, ctab); break;
    case COL_U32: lit_tile<uint32_t, NT, CJ_U>(q, tile, g0, ng, last, R, cov, red, ctab); break;
    case COL_U16: lit_tile<uint16_t, NT, CJ_U>(q, tile, g0, ng, last, R, cov, red, ctab); break;
    case COL_F32: lit_tile<float, NT, CJ_U>(q, tile, g0, ng, last, R, cov, red, ctab); break;
    default: lit_tile<double, NT, CJ_U>(q, tile, g0, ng, last, R, cov, red, ctab); break;
    }
}

// End of a (query, tile) work item of the batch kernels: the tile's result
// R[0, ng) goes to a.out (single-query launches), its ones to
// part[q*ntiles + tile], or, in the row-id pass, its row ids to a.rids.  A
// set bit b of group g is row 31g + 30 - b.  In the row-id pass thread i
// takes the PER consecutive groups from i*PER, so that a block-wide prefix
// sum of the threads' ones places every id.  red needs NT/64 + 1 words.
// Ends with a barrier.
template <int NT>
__device__ __forceinline__ void tile_epilogue(const ConjArgs& a, uint32_t qi, uint32_t tile, uint32_t g0,
                                              uint32_t ng, const word_t* R, uint32_t* red) {
    const int w = threadIdx.x / WAVE;
    if (a.rids) {
        constexpr int PER = (QT + NT - 1) / NT;
        const uint32_t k0 = threadIdx.x * (uint32_t)PER;
        uint32_t c = 0;
        #pragma unroll
        for (int i = 0; i < PER; ++i) c += k0 + i < ng ? (uint32_t)__popc(R[k0 + i]) : 0u;
        uint32_t tot;
        const uint32_t off = block_excl_sum<NT>(c, red, &tot);
        uint32_t* dst = a.rids + a.rbase[qi] + a.roff[(size_t)qi * a.ntiles + tile] + off;
        for (int i = 0; i < PER; ++i) {
            if (k0 + i >= ng) break;
            word_t v = R[k0 + i];
            const uint32_t r30 = (g0 + k0 + (uint32_t)i) * 31u + 30u;
            while (v) {
                const uint32_t b = 31u - (uint32_t)__clz(v);
                *dst++ = r30 - b;
                v &= ~(1u << b);
            }
        }
        __syncthreads();
        return;
    }
    uint32_t ones = 0;
    for (uint32_t i = threadIdx.x; i < ng; i += NT) {
        const word_t v = R[i];
        if (a.out) a.out[g0 + i] = v;
        ones += __popc(v);
    }
    ones = wave_sum(ones);
    if (lane_id() == 0) red[w] = ones;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < NT / WAVE; ++i) t += red[i];
        a.part[(size_t)qi * a.ntiles + tile] = t;
    }
    __syncthreads();
}

template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_conj(ConjArgs a) {
    __shared__ word_t R[QT + 1], R2[QT + 1];
    __shared__ uint32_t cov[QT + 1];
    __shared__ uint32_t red[NT / WAVE + 1];
    __shared__ uint64_t bal[NT / WAVE][32];
    __shared__ uint32_t ctab[CT_WORDS];
    const uint32_t nchunk = (a.ntiles + a.tpi - 1) / a.tpi;
    const uint32_t nitems = a.nq * nchunk;
    for (uint32_t k = threadIdx.x; k <= QT; k += NT) cov[k] = 0u;   // the literal plan keeps it zero
    __syncthreads();
    for (uint32_t item = blockIdx.x; item < nitems; item += gridDim.x) {
        const uint32_t qi = item / nchunk;
        const uint32_t c = item - qi * nchunk;
        const uint32_t k0 = a.qoff[qi], k1 = a.qoff[qi + 1];
        for (uint32_t tile = c * a.tpi; tile < min((c + 1) * a.tpi, a.ntiles); ++tile) {
            const bool last = tile + 1 == a.ntiles;
            const uint32_t g0 = tile * (uint32_t)QT;
            const uint32_t ng = last ? (a.groups - g0) + (a.anb ? 1u : 0u) : (uint32_t)QT;
            bool empty = false, counted = false;
            for (uint32_t k = k0; k < k1; ++k) {
                const QueryArgs q = a.terms[k];        // in registers (see k_range)
                term_tile<NT>(q, tile, g0, ng, last, k == k0 ? R : R2, cov, red, bal, ctab);
                // the last term of a count pass counts the ones as it ANDs
                const bool fuse = k + 1 == k1 && !a.rids && !a.out;
                int any = 0;
                uint32_t ones = 0;
                if (k != k0) {
                    for (uint32_t i = threadIdx.x; i < ng; i += NT) {
                        const word_t v = R[i] & R2[i];
                        R[i] = v;
                        any |= v != 0u;
                        ones += (uint32_t)__popc(v);
                    }
                } else {
                    for (uint32_t i = threadIdx.x; i < ng; i += NT) {
                        any |= R[i] != 0u;
                        ones += (uint32_t)__popc(R[i]);
                    }
                }
                if (fuse) {
                    // red is next written after a barrier that thread 0 passes
                    // only once it has read it
                    ones = wave_sum(ones);
                    if (lane_id() == 0) red[threadIdx.x / WAVE] = ones;
                    __syncthreads();
                    if (threadIdx.x == 0) {
                        uint32_t t = 0;
                        for (int i = 0; i < NT / WAVE; ++i) t += red[i];
                        a.part[(size_t)qi * a.ntiles + tile] = t;
                    }
                    counted = true;
                    break;
                }
                if (!__syncthreads_or(any)) {            // the tile is empty: skip the rest
                    empty = true;
                    break;
                }
            }
            if (counted) continue;
            if (empty && !a.out) {
                // no ones: the count is zero and the row-id pass has nothing
                // to write (the vote's barrier ended every read of R)
                if (!a.rids && threadIdx.x == 0) a.part[(size_t)qi * a.ntiles + tile] = 0u;
                continue;
            }
            // (an empty tile written out: R[0, ng) is all zero, as the vote found)
            tile_epilogue<NT>(a, qi, tile, g0, ng, R, red);
        }
    }
}

// Batches of single range terms (the range-query batch): the persistent
// form of k_query_lit over (query, TPI consecutive tiles) work items in
// grid-stride order, so that one launch serves the whole batch and the
// blocks of one query's last tiles overlap the next query's first.  The
// column type is a template parameter (one index per batch); SCAN 1 adds
// the raw-scan plan for batches that mix the plans (8 row loads in flight,
// registers shared with the literal plan), SCAN 2 is a batch of raw scans
// only (31 loads in flight).
// Counts only: each tile's ones go to part[q * ntiles + tile], the literal
// result is not written unless a.out is set (single-query batches).
template <class T, int NT, int SCAN>
__global__ __launch_bounds__(NT) void k_range(ConjArgs a) {
    __shared__ word_t R[QT + 1];
    __shared__ uint32_t cov[QT + 1];
    __shared__ uint32_t red[NT / WAVE + 1];
    __shared__ uint64_t bal[SCAN ? NT / WAVE : 1][32];
    __shared__ uint32_t ctab[CT_WORDS];
    const uint32_t nchunk = (a.ntiles + a.tpi - 1) / a.tpi;
    const uint32_t nitems = a.nq * nchunk;
    const int w = threadIdx.x / WAVE;
    for (uint32_t item = blockIdx.x; item < nitems; item += gridDim.x) {
        const uint32_t qi = item / nchunk;
        const uint32_t c = item - qi * nchunk;
        // a copy in registers: a reference into device memory would be
        // re-read after every store the compiler cannot tell apart from it
        const QueryArgs q = a.terms[a.qoff[qi]];
        for (uint32_t tile = c * a.tpi; tile < min((c + 1) * a.tpi, a.ntiles); ++tile) {
            const bool last = tile + 1 == a.ntiles;
            const uint32_t g0 = tile * (uint32_t)QT;
            const uint32_t ng = last ? (a.groups - g0) + (a.anb ? 1u : 0u) : (uint32_t)QT;
            if (SCAN == 2) {
                // raw scan only: each wave's groups assembled from ballots and
                // counted in registers (the structure of k_query_scan)
                const T* col = (const T*)q.col;
                const uint32_t nrows = q.groups * 31u + q.anb;
                const uint32_t lane = lane_id();
                uint32_t ones = 0;
                for (uint32_t k0 = (uint32_t)w * 64u; k0 < ng; k0 += (NT / WAVE) * 64u) {
                    scan_rows64<T, 31>(q, col, (g0 + k0) * 31u, nrows, bal[w]);
                    __builtin_amdgcn_wave_barrier();
                    const uint32_t k = k0 + lane;
                    if (k < ng) {
                        const uint32_t bit0 = lane * 31u;
                        const uint32_t kk = bit0 >> 6, s0 = bit0 & 63u;
                        uint64_t lo = bal[w][kk] >> s0;
                        if (s0 > 33u && kk + 1 < 31u) lo |= bal[w][kk + 1] << (64u - s0);
                        word_t v = __brev((uint32_t)(lo & 0x7FFFFFFFull)) >> 1;
                        if (g0 + k == q.groups) v &= (((1u << q.anb) - 1u) << (31u - q.anb));
                        if (a.out) a.out[g0 + k] = v;
                        ones += __popc(v);
                    }
                    __builtin_amdgcn_wave_barrier();
                }
                ones = wave_sum(ones);
                if (lane == 0) red[w] = ones;
                __syncthreads();
                if (threadIdx.x == 0) {
                    uint32_t t = 0;
                    for (int i = 0; i < NT / WAVE; ++i) t += red[i];
                    a.part[(size_t)qi * a.ntiles + tile] = t;
                }
                __syncthreads();
                continue;
            } else if (SCAN == 1 && q.plan == PLAN_SCAN_D) {
                scan_tile<T, NT>(q, g0, ng, R, bal);
            } else {
                for (uint32_t k = threadIdx.x; k <= QT; k += NT) { R[k] = 0u; cov[k] = 0u; }
                __syncthreads();
                lit_tile<T, NT, QB_U, false>(q, tile, g0, ng, last, R, cov, red, ctab);
            }
            tile_epilogue<NT>(a, qi, tile, g0, ng, R, red);
        }
    }
}

// Boolean expressions over range terms (countQuery::doEvaluate,
// src/countQuery.cpp:944-1003, with the mask of every row): per tile the
// block runs the query's postfix program over a stack of tile results in
// LDS.  NOT complements within the rows (an active group keeps only its
// nbits), NAND / NOR are NOT of AND / OR as fastbit_selection_combine
// builds them (src/iapi.cpp:1313-1326), MINUS is AND NOT.
template <int NT, int D, bool LIT>
__device__ __forceinline__ void expr_body(const ConjArgs& a) {
    __shared__ word_t S[D][QT + 1];      // D >= a.depth (the launcher's choice)
    __shared__ uint32_t cov[QT + 1];
    __shared__ uint32_t red[NT / WAVE + 1];
    __shared__ uint64_t bal[LIT ? 1 : NT / WAVE][32];
    __shared__ uint32_t ctab[CT_WORDS];
    const uint32_t nchunk = (a.ntiles + a.tpi - 1) / a.tpi;
    const uint32_t nitems = a.nq * nchunk;
    const word_t amask = a.anb ? (((1u << a.anb) - 1u) << (31u - a.anb)) : 0u;
    for (uint32_t k = threadIdx.x; k <= QT; k += NT) cov[k] = 0u;   // the literal plan keeps it zero
    __syncthreads();
    for (uint32_t item = blockIdx.x; item < nitems; item += gridDim.x) {
        const uint32_t qi = item / nchunk;
        const uint32_t c = item - qi * nchunk;
        const uint32_t p0 = a.poff[qi], p1 = a.poff[qi + 1];
        for (uint32_t tile = c * a.tpi; tile < min((c + 1) * a.tpi, a.ntiles); ++tile) {
            const bool last = tile + 1 == a.ntiles;
            const uint32_t g0 = tile * (uint32_t)QT;
            const uint32_t ng = last ? (a.groups - g0) + (a.anb ? 1u : 0u) : (uint32_t)QT;
            int sp = 0;
            bool empty = false;
            for (uint32_t p = p0; p < p1; ++p) {
                const int32_t sk = a.skip[p];
                if (sk >= 0) {
                    int any = 0;
                    for (uint32_t i = threadIdx.x; i < ng; i += NT) any |= S[sp - 1][i] != 0u;
                    if (!__syncthreads_or(any)) {
                        if ((uint32_t)sk == p1 && sp == 1) {   // the empty operand is the result
                            empty = true;
                            break;
                        }
                        p = (uint32_t)sk - 1u;          // the loop's ++p resumes at sk
                        continue;
                    }
                }
                const int32_t op = a.prog[p];
                if (op >= 0) {
                    const QueryArgs q = a.terms[op];       // in registers (see k_range)
                    term_tile<NT, LIT>(q, tile, g0, ng, last, S[sp], cov, red, bal, ctab);
                    ++sp;
                    continue;
                }
                if (op == EXPR_NOT) {
                    word_t* x = S[sp - 1];
                    for (uint32_t i = threadIdx.x; i < ng; i += NT)
                        x[i] = ~x[i] & (g0 + i == a.groups ? amask : 0x7FFFFFFFu);
                } else {
                    word_t* x = S[sp - 2];
                    const word_t* y = S[sp - 1];
                    for (uint32_t i = threadIdx.x; i < ng; i += NT) {
                        const word_t u = x[i], v = y[i];
                        const word_t m = g0 + i == a.groups ? amask : 0x7FFFFFFFu;
                        word_t r;
                        switch (op) {
                        case EXPR_AND: r = u & v; break;
                        case EXPR_OR: r = u | v; break;
                        case EXPR_XOR: r = u ^ v; break;
                        case EXPR_NAND: r = ~(u & v) & m; break;
                        case EXPR_NOR: r = ~(u | v) & m; break;
                        default: r = u & ~v; break;          // EXPR_MINUS
                        }
                        x[i] = r;
                    }
                    --sp;
                }
                __syncthreads();
            }
            if (empty && !a.out) {
                // no ones: the count is zero and the row-id pass has nothing
                // to write (the vote's barrier ended every read of S)
                if (!a.rids && threadIdx.x == 0) a.part[(size_t)qi * a.ntiles + tile] = 0u;
                continue;
            }
            tile_epilogue<NT>(a, qi, tile, g0, ng, S[0], red);
        }
    }
}

template <int NT, int D>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_expr(ConjArgs a) { expr_body<NT, D, false>(a); }

// Programs of literal terms with no edge bitmaps (relic, fuzz, zona): the
// scan and edge-check code left out, registers held to 6 waves per SIMD.
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_expr_lit(ConjArgs a) {
    expr_body<NT, 2, true>(a);
}


// Raw scan plan: R[g] = rows 31g..31g+30 that satisfy the range.  A wave
// covers 64 groups (1984 rows): 31 coalesced loads of 64 values, 31
// ballots staged in LDS, each lane assembles its group.
template <class T>
__global__ __launch_bounds__(IX_TPB) void k_query_scan(QueryArgs q, uint32_t nrows) {
    __shared__ uint64_t bal[IX_TPB / WAVE][32];
    const int lane = (int)lane_id();
    const int w = threadIdx.x / WAVE;
    const T* col = (const T*)q.col;
    const uint32_t ngroups_all = q.groups + (q.anb ? 1u : 0u);
    uint32_t ones = 0;
    for (uint32_t gb = (blockIdx.x * (IX_TPB / WAVE) + w) * 64u; gb < ngroups_all;
         gb += gridDim.x * (IX_TPB / WAVE) * 64u) {
        const uint32_t rbase = gb * 31u;
        scan_rows64<T, 31>(q, col, rbase, nrows, bal[w]);
        __builtin_amdgcn_wave_barrier();
        const uint32_t g = gb + lane;
        if (g < ngroups_all) {
            const uint32_t bit0 = (uint32_t)lane * 31u;      // in the 1984-bit stream
            const uint32_t k0 = bit0 >> 6, s0 = bit0 & 63u;
            uint64_t lo = bal[w][k0] >> s0;
            if (s0 > 33u && k0 + 1 < 31u) lo |= bal[w][k0 + 1] << (64u - s0);
            const uint32_t bits = (uint32_t)(lo & 0x7FFFFFFFull);   // bit j = row 31g+j
            word_t v = __brev(bits) >> 1;                          // row 31g -> bit 30
            if (g == q.groups) v &= (((1u << q.anb) - 1u) << (31u - q.anb));
            q.out[g] = v;
            ones += __popc(v);
        }
        __builtin_amdgcn_wave_barrier();
    }
    __shared__ uint32_t red[IX_TPB / WAVE];
    ones = wave_sum(ones);
    if (lane == 0) red[w] = ones;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < IX_TPB / WAVE; ++i) t += red[i];
        q.part[blockIdx.x] = t;
    }
}

// counts[i] = sum of parts[i * stride .. i * stride + np[i]), one block per
// query (np == nullptr: nfixed parts for every query)
__global__ __launch_bounds__(IX_TPB) void k_sum_parts(const uint32_t* __restrict__ parts, uint32_t stride,
                                                      const uint32_t* __restrict__ np, uint32_t nfixed,
                                                      uint32_t* __restrict__ counts) {
    __shared__ uint32_t red[IX_TPB / WAVE];
    const uint32_t qi = blockIdx.x;
    const uint32_t n = np ? np[qi] : nfixed;
    uint32_t t = 0;
    for (uint32_t k = threadIdx.x; k < n; k += IX_TPB) t += parts[(size_t)qi * stride + k];
    t = wave_sum(t);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t x = 0;
        for (int i = 0; i < IX_TPB / WAVE; ++i) x += red[i];
        counts[qi] = x;
    }
}

// roff[q*stride + t] = parts[q*stride] + ... + parts[q*stride + t - 1] for
// t < n, one block per query, IX_TPB entries per round.
__global__ __launch_bounds__(IX_TPB) void k_scan_parts(const uint32_t* __restrict__ parts, uint32_t stride,
                                                       uint32_t n, uint32_t* __restrict__ roff) {
    __shared__ uint32_t sh[IX_TPB / WAVE + 1];
    const size_t base = (size_t)blockIdx.x * stride;
    uint32_t carry = 0;
    for (uint32_t k0 = 0; k0 < n; k0 += IX_TPB) {
        const uint32_t k = k0 + threadIdx.x;
        const uint32_t v = k < n ? parts[base + k] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_sum<IX_TPB>(v, sh, &tot);
        if (k < n) roff[base + k] = carry + ex;
        carry += tot;
    }
}

// ---------------------------------------------------------------------------
// Distinct values and bitmap starts.
// Order-preserving 32-bit key of a value.
template <class T> __device__ __forceinline__ uint32_t order_key(T v);
template <> __device__ __forceinline__ uint32_t order_key<int32_t>(int32_t v) { return (uint32_t)v ^ 0x80000000u; }
template <> __device__ __forceinline__ uint32_t order_key<uint32_t>(uint32_t v) { return v; }
template <> __device__ __forceinline__ uint32_t order_key<uint16_t>(uint16_t v) { return v; }
template <> __device__ __forceinline__ uint32_t order_key<float>(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

template <class T>
__global__ __launch_bounds__(IX_TPB) void k_order_keys(const T* __restrict__ col, uint32_t n,
                                                       uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * IX_TPB + threadIdx.x;
    if (i < n) keys[i] = order_key<T>(col[i]);
}

// flag[i] = 1 when sorted key i differs from key i-1
__global__ __launch_bounds__(IX_TPB) void k_unique_flags(const uint32_t* __restrict__ k, uint32_t n,
                                                         uint32_t* __restrict__ flag) {
    const uint32_t i = blockIdx.x * IX_TPB + threadIdx.x;
    if (i < n) flag[i] = (i == 0 || k[i] != k[i - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(IX_TPB) void k_unique_scatter(const uint32_t* __restrict__ k, uint32_t n,
                                                           const uint32_t* __restrict__ flag,
                                                           const uint32_t* __restrict__ pos,
                                                           uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * IX_TPB + threadIdx.x;
    if (i < n && flag[i]) out[pos[i]] = k[i];
}

// first[pos[i]] = i for every i that starts a run of equal sorted keys
__global__ __launch_bounds__(IX_TPB) void k_unique_first(const uint32_t* __restrict__ flag,
                                                         const uint32_t* __restrict__ pos, uint32_t n,
                                                         uint32_t* __restrict__ first) {
    const uint32_t i = blockIdx.x * IX_TPB + threadIdx.x;
    if (i < n && flag[i]) first[pos[i]] = i;
}

// start[b] = first index of key b in the sorted keys (b in [0, nbm])
__global__ __launch_bounds__(IX_TPB) void k_starts(const uint32_t* __restrict__ k, uint32_t n,
                                                   uint32_t nbm, uint32_t* __restrict__ start) {
    const uint32_t b = blockIdx.x * IX_TPB + threadIdx.x;
    if (b > nbm) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (k[mid] < b) lo = mid + 1; else hi = mid;
    }
    start[b] = lo;
}

// ---------------------------------------------------------------------------
// Host launchers.
static inline uint32_t cdivu(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

hipError_t launch_keys_relic(const void* col, int type, uint32_t n, const double* vals, uint32_t nv,
                             uint32_t* keys, hipStream_t s) {
    dim3 g(cdivu(n, IX_TPB)), b(IX_TPB);
    switch (type) {
    case COL_I32: hipLaunchKernelGGL(k_keys_relic<int32_t>, g, b, 0, s, (const int32_t*)col, n, vals, nv, keys); break;
    case COL_U32: hipLaunchKernelGGL(k_keys_relic<uint32_t>, g, b, 0, s, (const uint32_t*)col, n, vals, nv, keys); break;
    case COL_U16: hipLaunchKernelGGL(k_keys_relic<uint16_t>, g, b, 0, s, (const uint16_t*)col, n, vals, nv, keys); break;
    case COL_F32: hipLaunchKernelGGL(k_keys_relic<float>, g, b, 0, s, (const float*)col, n, vals, nv, keys); break;
    case COL_F64: hipLaunchKernelGGL(k_keys_relic<double>, g, b, 0, s, (const double*)col, n, vals, nv, keys); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_keys_card(const uint32_t* ind, uint32_t n, uint32_t card, uint32_t* keys,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_keys_card, dim3(cdivu(n, IX_TPB)), dim3(IX_TPB), 0, s, ind, n, card, keys);
    return hipGetLastError();
}

hipError_t launch_keys_bin(const void* col, int type, uint32_t n, const double* bounds, uint32_t nobs,
                           uint32_t* keys, hipStream_t s) {
    dim3 g(cdivu(n, IX_TPB)), b(IX_TPB);
    switch (type) {
    case COL_I32: hipLaunchKernelGGL(k_keys_bin<int32_t>, g, b, 0, s, (const int32_t*)col, n, bounds, nobs, keys); break;
    case COL_U32: hipLaunchKernelGGL(k_keys_bin<uint32_t>, g, b, 0, s, (const uint32_t*)col, n, bounds, nobs, keys); break;
    case COL_U16: hipLaunchKernelGGL(k_keys_bin<uint16_t>, g, b, 0, s, (const uint16_t*)col, n, bounds, nobs, keys); break;
    case COL_F32: hipLaunchKernelGGL(k_keys_bin<float>, g, b, 0, s, (const float*)col, n, bounds, nobs, keys); break;
    case COL_F64: hipLaunchKernelGGL(k_keys_bin<double>, g, b, 0, s, (const double*)col, n, bounds, nobs, keys); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_order_keys(const void* col, int type, uint32_t n, uint32_t* keys, hipStream_t s) {
    dim3 g(cdivu(n, IX_TPB)), b(IX_TPB);
    switch (type) {
    case COL_I32: hipLaunchKernelGGL(k_order_keys<int32_t>, g, b, 0, s, (const int32_t*)col, n, keys); break;
    case COL_U32: hipLaunchKernelGGL(k_order_keys<uint32_t>, g, b, 0, s, (const uint32_t*)col, n, keys); break;
    case COL_U16: hipLaunchKernelGGL(k_order_keys<uint16_t>, g, b, 0, s, (const uint16_t*)col, n, keys); break;
    case COL_F32: hipLaunchKernelGGL(k_order_keys<float>, g, b, 0, s, (const float*)col, n, keys); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Distinct sorted keys: out receives them, *ctl the count (device).
// tmp needs 2*n + scan_tmp_words(n) words.
hipError_t launch_unique(const uint32_t* k, uint32_t n, uint32_t* out, uint32_t* tmp, hipStream_t s) {
    if (!n) return hipSuccess;
    uint32_t* flag = tmp;
    uint32_t* pos = tmp + n;
    hipLaunchKernelGGL(k_unique_flags, dim3(cdivu(n, IX_TPB)), dim3(IX_TPB), 0, s, k, n, flag);
    hipError_t e = launch_scan_excl(flag, pos, n, tmp + 2 * n, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_unique_scatter, dim3(cdivu(n, IX_TPB)), dim3(IX_TPB), 0, s, k, n,
                       (const uint32_t*)flag, (const uint32_t*)pos, out);
    return hipGetLastError();
}

hipError_t launch_unique_first(const uint32_t* flag, const uint32_t* pos, uint32_t n, uint32_t* first,
                               hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_unique_first, dim3(cdivu(n, IX_TPB)), dim3(IX_TPB), 0, s, flag, pos, n, first);
    return hipGetLastError();
}

hipError_t launch_starts(const uint32_t* k, uint32_t n, uint32_t nbm, uint32_t* start, hipStream_t s) {
    hipLaunchKernelGGL(k_starts, dim3(cdivu(nbm + 1, IX_TPB)), dim3(IX_TPB), 0, s, k, n, nbm, start);
    return hipGetLastError();
}

uint32_t rs_blocks(uint32_t n) { return n ? cdivu(n, RS_TILE) : 1u; }

// One radix pass: (ki, vi) -> (ko, vo) by digit at `shift`.  tmp needs
// 2*256*rs_blocks(n) + scan_tmp_words(256*rs_blocks(n)) words.
hipError_t launch_rs_pass(const uint32_t* ki, const uint32_t* vi, uint32_t n, uint32_t shift,
                          uint32_t* ko, uint32_t* vo, uint32_t* tmp, hipStream_t s) {
    const uint32_t nb = rs_blocks(n);
    uint32_t* hist = tmp;
    uint32_t* offs = tmp + 256u * nb;
    hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_TPB), 0, s, ki, n, shift, nb, hist);
    hipError_t e = launch_scan_excl(hist, offs, 256u * nb, tmp + 512u * nb, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_TPB), 0, s, ki, vi, n, shift, nb,
                       (const uint32_t*)offs, ko, vo);
    return hipGetLastError();
}

template <bool WRITE>
static hipError_t launch_enc_t(const EncArgs& a, int type, const void* col, hipStream_t s) {
    if (!a.nchunk) return hipSuccess;
    dim3 g(min(a.nchunk, 1u << 16)), b(IX_TPB);
    if (!col) {
        hipLaunchKernelGGL((k_enc_chunks<float, WRITE>), g, b, 0, s, a, (const float*)nullptr);
        return hipGetLastError();
    }
    switch (type) {
    case COL_I32: hipLaunchKernelGGL((k_enc_chunks<int32_t, WRITE>), g, b, 0, s, a, (const int32_t*)col); break;
    case COL_U32: hipLaunchKernelGGL((k_enc_chunks<uint32_t, WRITE>), g, b, 0, s, a, (const uint32_t*)col); break;
    case COL_U16: hipLaunchKernelGGL((k_enc_chunks<uint16_t, WRITE>), g, b, 0, s, a, (const uint16_t*)col); break;
    case COL_F32: hipLaunchKernelGGL((k_enc_chunks<float, WRITE>), g, b, 0, s, a, (const float*)col); break;
    case COL_F64: hipLaunchKernelGGL((k_enc_chunks<double, WRITE>), g, b, 0, s, a, (const double*)col); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_enc_count(const EncArgs& a, int type, const void* col, hipStream_t s) {
    return launch_enc_t<false>(a, type, col, s);
}

hipError_t launch_enc_write(const EncArgs& a, hipStream_t s) {
    return launch_enc_t<true>(a, COL_F32, nullptr, s);
}

hipError_t launch_toff_len(const uint32_t* words, const uint64_t* off, const uint32_t* nw, const uint32_t* tb,
                           const uint32_t* tw0, uint32_t nch, uint32_t* clen, hipStream_t s) {
    if (!nch) return hipSuccess;
    hipLaunchKernelGGL(k_toff_len, dim3(min(nch, 1u << 16)), dim3(IX_TPB), 0, s, words, off, nw, tb, tw0, nch, clen);
    return hipGetLastError();
}

hipError_t launch_toff_rows(const uint32_t* words, const uint64_t* off, const uint32_t* nw, const uint32_t* tb,
                            const uint32_t* tw0, const uint32_t* cg0, uint32_t nch, uint32_t nbm,
                            uint32_t ntiles, uint32_t tg, uint2* toff, hipStream_t s) {
    if (!nch) return hipSuccess;
    hipLaunchKernelGGL(k_toff_rows, dim3(min(nch, 1u << 16)), dim3(IX_TPB), 0, s, words, off, nw, tb, tw0, cg0,
                       nch, nbm, ntiles, tg, toff);
    return hipGetLastError();
}

hipError_t launch_toff_tail(const uint64_t* off, const uint32_t* nw, const uint32_t* total, uint32_t nbm,
                            uint32_t ntiles, uint32_t tg, uint32_t groups, uint2* toff, hipStream_t s) {
    if (!nbm) return hipSuccess;
    hipLaunchKernelGGL(k_toff_tail, dim3(cdivu(nbm, IX_TPB)), dim3(IX_TPB), 0, s, off, nw, total, nbm, ntiles,
                       tg, groups, toff);
    return hipGetLastError();
}

hipError_t launch_tseg_len(const uint2* toff, uint32_t nobs, uint32_t ntiles, uint32_t groups,
                           uint32_t anb, uint32_t* len, hipStream_t s) {
    const uint32_t n = ntiles * nobs + 1;
    hipLaunchKernelGGL(k_tseg_len, dim3(cdivu(n, IX_TPB)), dim3(IX_TPB), 0, s, toff, nobs, ntiles, groups,
                       anb, len);
    return hipGetLastError();
}

hipError_t launch_tseg_fill(const uint32_t* words, const uint2* toff, const uint32_t* TW, uint32_t nobs,
                            uint32_t ntiles, uint32_t groups, uint32_t anb, uint32_t* tw, hipStream_t s) {
    const uint32_t n = ntiles * nobs;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_tseg_fill, dim3(min(cdivu(n, IX_TPB / WAVE), 1u << 20)), dim3(IX_TPB), 0, s, words,
                       toff, TW, nobs, ntiles, groups, anb, tw);
    return hipGetLastError();
}

uint32_t scan_blocks(uint32_t groups, uint32_t anb) {
    const uint32_t waves = cdivu(groups + (anb ? 1u : 0u), 64u);
    const uint32_t blocks = min(cdivu(waves, IX_TPB / WAVE), (uint32_t)SCAN_MAX_BLOCKS);
    return blocks ? blocks : 1u;
}

hipError_t launch_sum_parts(const uint32_t* parts, uint32_t stride, const uint32_t* np, uint32_t nq,
                            uint32_t* counts, hipStream_t s) {
    if (!nq) return hipSuccess;
    hipLaunchKernelGGL(k_sum_parts, dim3(nq), dim3(IX_TPB), 0, s, parts, stride, np, 0u, counts);
    return hipGetLastError();
}

hipError_t launch_sum_parts_n(const uint32_t* parts, uint32_t n, uint32_t* count, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(IX_TPB), 0, s, parts, n, (const uint32_t*)nullptr, n,
                       count);
    return hipGetLastError();
}

hipError_t launch_scan_parts(const uint32_t* parts, uint32_t stride, uint32_t n, uint32_t nq, uint32_t* roff,
                             hipStream_t s) {
    if (!nq || !n) return hipSuccess;
    hipLaunchKernelGGL(k_scan_parts, dim3(nq), dim3(IX_TPB), 0, s, parts, stride, n, roff);
    return hipGetLastError();
}

template <int NT>
static hipError_t launch_query_lit_nt(const QueryArgs& q, int type, hipStream_t s) {
    dim3 g(q.ntiles), b(NT);                 // the last tile holds the active slot
    switch (type) {
    case COL_I32: hipLaunchKernelGGL((k_query_lit<int32_t, NT>), g, b, 0, s, q); break;
    case COL_U32: hipLaunchKernelGGL((k_query_lit<uint32_t, NT>), g, b, 0, s, q); break;
    case COL_U16: hipLaunchKernelGGL((k_query_lit<uint16_t, NT>), g, b, 0, s, q); break;
    case COL_F32: hipLaunchKernelGGL((k_query_lit<float, NT>), g, b, 0, s, q); break;
    case COL_F64: hipLaunchKernelGGL((k_query_lit<double, NT>), g, b, 0, s, q); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Block size of the literal plan: 256 threads (8 blocks per CU) while all
// tiles fit on the device at once, else 192 (9 blocks per CU by LDS), so
// that, e.g., the 2114 tiles of 2^27 rows run in one wave of blocks on 256
// CUs instead of 2048 blocks and a tail of 66.
hipError_t launch_query_lit(const QueryArgs& q, int type, hipStream_t s) {
    static const uint32_t cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        return (uint32_t)n;
    }();
    if (q.ntiles <= 8u * cus) return launch_query_lit_nt<256>(q, type, s);
    return launch_query_lit_nt<192>(q, type, s);
}

hipError_t launch_conj(const ConjArgs& a, int nt, uint32_t blocks, hipStream_t s) {
    if (!a.nq) return hipSuccess;
    if (a.prog) {
        // the stack in LDS only as deep as the batch's programs need: a
        // shallower stack leaves room for more blocks per CU
        if (a.depth == 0 || a.depth > (uint32_t)EXPR_DEPTH) return hipErrorInvalidValue;
        const bool d2 = a.depth <= 2;
        if (a.lit_only && d2) {
            switch (nt) {
            case 128: hipLaunchKernelGGL((k_expr_lit<128>), dim3(blocks), dim3(128), 0, s, a); break;
            case 256: hipLaunchKernelGGL((k_expr_lit<256>), dim3(blocks), dim3(256), 0, s, a); break;
            case 512: hipLaunchKernelGGL((k_expr_lit<512>), dim3(blocks), dim3(512), 0, s, a); break;
            default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        }
        switch (nt) {
        case 128:
            if (d2) hipLaunchKernelGGL((k_expr<128, 2>), dim3(blocks), dim3(128), 0, s, a);
            else hipLaunchKernelGGL((k_expr<128, EXPR_DEPTH>), dim3(blocks), dim3(128), 0, s, a);
            break;
        case 256:
            if (d2) hipLaunchKernelGGL((k_expr<256, 2>), dim3(blocks), dim3(256), 0, s, a);
            else hipLaunchKernelGGL((k_expr<256, EXPR_DEPTH>), dim3(blocks), dim3(256), 0, s, a);
            break;
        case 512:
            if (d2) hipLaunchKernelGGL((k_expr<512, 2>), dim3(blocks), dim3(512), 0, s, a);
            else hipLaunchKernelGGL((k_expr<512, EXPR_DEPTH>), dim3(blocks), dim3(512), 0, s, a);
            break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (nt) {
    case 128: hipLaunchKernelGGL(k_conj<128>, dim3(blocks), dim3(128), 0, s, a); break;
    case 256: hipLaunchKernelGGL(k_conj<256>, dim3(blocks), dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL(k_conj<512>, dim3(blocks), dim3(512), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class T, int SCAN>
static hipError_t launch_range_t(const ConjArgs& a, int nt, uint32_t blocks, hipStream_t s) {
    switch (nt) {
    case 192: hipLaunchKernelGGL((k_range<T, 192, SCAN>), dim3(blocks), dim3(192), 0, s, a); break;
    case 256: hipLaunchKernelGGL((k_range<T, 256, SCAN>), dim3(blocks), dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int SCAN>
static hipError_t launch_range_s(const ConjArgs& a, int type, int nt, uint32_t blocks, hipStream_t s) {
    switch (type) {
    case COL_I32: return launch_range_t<int32_t, SCAN>(a, nt, blocks, s);
    case COL_U32: return launch_range_t<uint32_t, SCAN>(a, nt, blocks, s);
    case COL_U16: return launch_range_t<uint16_t, SCAN>(a, nt, blocks, s);
    case COL_F32: return launch_range_t<float, SCAN>(a, nt, blocks, s);
    case COL_F64: return launch_range_t<double, SCAN>(a, nt, blocks, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_range(const ConjArgs& a, int type, int scan, int nt, uint32_t blocks, hipStream_t s) {
    if (!a.nq) return hipSuccess;
    switch (scan) {
    case 0: return launch_range_s<0>(a, type, nt, blocks, s);
    case 1: return launch_range_s<1>(a, type, nt, blocks, s);
    default: return launch_range_s<2>(a, type, nt, blocks, s);
    }
}

__global__ void k_wah_edges(const uint32_t* out, uint32_t* ctl) {
    if (threadIdx.x == 0) {
        const uint32_t n = ctl[1];
        ctl[4] = n ? out[0] : 0u;
        ctl[5] = n ? out[n - 1] : 0u;
    }
}

uint32_t wah_status_words(uint32_t ntiles) { return ntiles + ntiles / 64u + 1u; }

hipError_t launch_wah_edges(const uint32_t* out, uint32_t* ctl, hipStream_t s) {
    hipLaunchKernelGGL(k_wah_edges, dim3(1), dim3(64), 0, s, out, ctl);
    return hipGetLastError();
}

hipError_t launch_query_wah(const QueryArgs& q, int type, const WahOut& o, hipStream_t s) {
    if (o.tag != 1u && o.tag != 2u) return hipErrorInvalidValue;
    dim3 g(q.ntiles), b(256);
    switch (type) {
    case COL_I32: hipLaunchKernelGGL((k_query_wah<int32_t, 256>), g, b, 0, s, q, o); break;
    case COL_U32: hipLaunchKernelGGL((k_query_wah<uint32_t, 256>), g, b, 0, s, q, o); break;
    case COL_U16: hipLaunchKernelGGL((k_query_wah<uint16_t, 256>), g, b, 0, s, q, o); break;
    case COL_F32: hipLaunchKernelGGL((k_query_wah<float, 256>), g, b, 0, s, q, o); break;
    case COL_F64: hipLaunchKernelGGL((k_query_wah<double, 256>), g, b, 0, s, q, o); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_query_scan(const QueryArgs& q, int type, uint32_t nrows, hipStream_t s) {
    dim3 g(scan_blocks(q.groups, q.anb)), b(IX_TPB);
    switch (type) {
    case COL_I32: hipLaunchKernelGGL(k_query_scan<int32_t>, g, b, 0, s, q, nrows); break;
    case COL_U32: hipLaunchKernelGGL(k_query_scan<uint32_t>, g, b, 0, s, q, nrows); break;
    case COL_U16: hipLaunchKernelGGL(k_query_scan<uint16_t>, g, b, 0, s, q, nrows); break;
    case COL_F32: hipLaunchKernelGGL(k_query_scan<float>, g, b, 0, s, q, nrows); break;
    case COL_F64: hipLaunchKernelGGL(k_query_scan<double>, g, b, 0, s, q, nrows); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace fbk
