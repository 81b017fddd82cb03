// Shared definitions of the index kernels and their host launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbk {

constexpr uint32_t NOKEY = 0xFFFFFFFFu;
constexpr double DBL_MAX_D = 1.7976931348623157e308;
constexpr int QT = 2048;                 // groups per query tile (LDS words)
constexpr int SCAN_MAX_BLOCKS = 2048;    // blocks of the raw scan plan

// column types
enum : int { COL_I32 = 0, COL_U32 = 1, COL_U16 = 2, COL_F32 = 3, COL_F64 = 4 };
// comparison operators of one side of a range
enum : int { OPC_NONE = 0, OPC_LT = 1, OPC_LE = 2 };
// plans of a term inside a conjunction
enum : int { PLAN_LIT_D = 1, PLAN_SCAN_D = 3 };

struct QueryArgs {
    const uint32_t* tw;        // tiled copy of the bitmaps
    const uint32_t* TW;        // segment (t, b) at tw[TW[t*nobs + b]], ntiles*nobs + 1 entries
    uint32_t ntiles;
    uint32_t nobs;
    uint32_t groups;           // whole 31-bit groups per bitmap
    uint32_t anb;              // active bits per bitmap
    uint32_t r0, r1, r2, r3;   // OR bitmaps [r0,r1) and [r2,r3)
    uint32_t flip;             // complement the OR (complement strategy)
    uint32_t cand0, cand1;     // edge bitmaps to check against the column
    const void* col;           // raw column
    int is_int;
    int lop, rop;
    double lo, hi;
    int64_t ilo, ihi;
    uint32_t* out;             // literal result, groups + 1 words
    uint32_t* part;            // ones per block (ntiles or scan_blocks entries)
    int type;                  // column type (conjunction terms)
    int plan;                  // PLAN_LIT_D or PLAN_SCAN_D (conjunction terms)
};

// Output of the compressed plan: the result's code words and chunk index
// (start group of every 32nd word, closed by the group count), the tiles'
// run summaries (prefix_summary_seg: ntiles aggregates, then ntiles/64 + 1
// segment aggregates; zeroed when allocated, tag alternating 1 / 2 between
// launches on the same buffers), and ctl: [1] the word count, [2] the active
// group's value (left-aligned bits).
struct WahOut {
    uint32_t* out;
    uint32_t* outck;
    uint64_t* status;
    uint32_t* ctl;
    uint32_t tag;
};

// A batch of conjunctions: query q ANDs terms[qoff[q] .. qoff[q+1]); all
// terms are over columns of the same rows.  part[q*ntiles + t] receives the
// ones of tile t; out (may be null) the literal result of a single query.
struct ConjArgs {
    const QueryArgs* terms;
    const uint32_t* qoff;
    uint32_t nq;
    uint32_t ntiles;
    uint32_t groups, anb;
    uint32_t tpi;              // tiles per work item
    uint32_t* part;
    uint32_t* out;
    // boolean expressions instead of conjunctions when prog is not null:
    // query q runs the postfix program prog[poff[q] .. poff[q+1]); an item
    // >= 0 pushes the result of terms[item], an EXPR_* item combines the
    // top of the stack (at most EXPR_DEPTH results)
    const int32_t* prog;
    const uint32_t* poff;
    uint32_t depth;            // deepest stack any program reaches
    // skip[p] >= 0: item p starts the right operand of an AND or MINUS;
    // when the left operand's tile is empty the program resumes at skip[p]
    // (the empty left result is the result of the operator)
    const int32_t* skip;
    // row-id pass (countQuery::getHitRows): when rids is set, tile t of
    // query q writes its row ids, ascending, from
    // rids[rbase[q] + roff[q*ntiles + t]] on, and part is not written
    uint32_t* rids;
    const uint64_t* rbase;
    const uint32_t* roff;
    // every term runs the literal plan with no edge bitmaps (relic and
    // two-level terms): the expression kernel takes its lighter form
    uint32_t lit_only;
};

constexpr int EXPR_DEPTH = 4;
constexpr int32_t EXPR_AND = -1, EXPR_OR = -2, EXPR_XOR = -3, EXPR_NAND = -4, EXPR_NOR = -5,
                  EXPR_NOT = -6, EXPR_MINUS = -7;

hipError_t launch_keys_relic(const void* col, int type, uint32_t n, const double* vals, uint32_t nv,
                             uint32_t* keys, hipStream_t s);
hipError_t launch_keys_card(const uint32_t* ind, uint32_t n, uint32_t card, uint32_t* keys,
                            hipStream_t s);
hipError_t launch_keys_bin(const void* col, int type, uint32_t n, const double* bounds, uint32_t nobs,
                           uint32_t* keys, hipStream_t s);
hipError_t launch_order_keys(const void* col, int type, uint32_t n, uint32_t* keys, hipStream_t s);
hipError_t launch_unique(const uint32_t* k, uint32_t n, uint32_t* out, uint32_t* tmp, hipStream_t s);
// after launch_unique: first[j] = sorted position of distinct key j
// (flag = tmp, pos = tmp + n of that call)
hipError_t launch_unique_first(const uint32_t* flag, const uint32_t* pos, uint32_t n, uint32_t* first,
                               hipStream_t s);
hipError_t launch_starts(const uint32_t* k, uint32_t n, uint32_t nbm, uint32_t* start, hipStream_t s);
uint32_t rs_blocks(uint32_t n);
hipError_t launch_rs_pass(const uint32_t* ki, const uint32_t* vi, uint32_t n, uint32_t shift,
                          uint32_t* ko, uint32_t* vo, uint32_t* tmp, hipStream_t s);
// Chunk-parallel bitmap encoding (k_enc_chunks): chunk c is rows
// [cstart[c], min(cstart[c] + ENC_CR, start[b+1])) of bitmap cbm[c].
constexpr uint32_t ENC_ROWS = 2048;          // rows per chunk (ENC_CR)
struct EncArgs {
    const uint32_t* rows;       // row ids sorted by bitmap, ascending within one
    const uint32_t* start;      // nbm + 1 bitmap starts
    const uint32_t* cbm;        // chunk -> bitmap
    const uint32_t* cstart;     // chunk -> first sorted position
    uint32_t nchunk;
    uint32_t groups, anb;
    // count pass
    uint32_t* ccnt;             // code words per chunk
    uint32_t* aval;             // active bits per bitmap (zeroed before)
    double* cmin;               // value range per chunk (bins)
    double* cmax;
    // write pass
    uint32_t* out;              // the serialized bitmaps
    const uint64_t* wpos;       // chunk's first word in out (~0: skip the chunk)
    const uint64_t* boff;       // bitmap's first word in out
    const uint32_t* bnw;        // bitmap's words before its trailer
    const uint8_t* unpack;      // bitmap stored decompressed (null: none)
};
hipError_t launch_enc_count(const EncArgs& a, int type, const void* col, hipStream_t s);
hipError_t launch_enc_write(const EncArgs& a, hipStream_t s);
// Tile offset table (ntiles + 1 rows of nbm entries): row t < ntiles holds,
// per bitmap, {absolute index of the first code word whose span ends after
// group t*tg, its start group} ({end, groups} past the last word); row
// ntiles holds {end of the code words, 1 when the bitmap is not null}.
// Built over chunks of TOFF_WORDS code words (chunk c: words
// [tw0[c], tw0[c] + TOFF_WORDS) of bitmap tb[c]): pass 1 gives each
// chunk's groups, the host scans them per bitmap into cg0 (and total[b]),
// pass 2 writes the rows, the tail pass the rest.
constexpr uint32_t TOFF_WORDS = 4096;
hipError_t launch_toff_len(const uint32_t* words, const uint64_t* off, const uint32_t* nw, const uint32_t* tb,
                           const uint32_t* tw0, uint32_t nch, uint32_t* clen, hipStream_t s);
hipError_t launch_toff_rows(const uint32_t* words, const uint64_t* off, const uint32_t* nw, const uint32_t* tb,
                            const uint32_t* tw0, const uint32_t* cg0, uint32_t nch, uint32_t nbm,
                            uint32_t ntiles, uint32_t tg, uint2* toff, hipStream_t s);
hipError_t launch_toff_tail(const uint64_t* off, const uint32_t* nw, const uint32_t* total, uint32_t nbm,
                            uint32_t ntiles, uint32_t tg, uint32_t groups, uint2* toff, hipStream_t s);
// Tiled copy: len[t*nobs + b] = words of segment (t, b) (ntiles*nobs + 1
// entries, the last 0; its exclusive scan is TW), then the copy itself.
hipError_t launch_tseg_len(const uint2* toff, uint32_t nobs, uint32_t ntiles, uint32_t groups,
                           uint32_t anb, uint32_t* len, hipStream_t s);
hipError_t launch_tseg_fill(const uint32_t* words, const uint2* toff, const uint32_t* TW, uint32_t nobs,
                            uint32_t ntiles, uint32_t groups, uint32_t anb, uint32_t* tw, hipStream_t s);
hipError_t launch_query_lit(const QueryArgs& q, int type, hipStream_t s);
hipError_t launch_query_scan(const QueryArgs& q, int type, uint32_t nrows, hipStream_t s);
// The compressed plan, one launch of one block per tile: o.out needs
// q.groups + 1 words, o.outck q.groups / 32 + 2 entries, o.status
// wah_status_words(q.ntiles) entries, o.ctl 8 words.
uint32_t wah_status_words(uint32_t ntiles);
hipError_t launch_query_wah(const QueryArgs& q, int type, const WahOut& o, hipStream_t s);
// ctl[4] / ctl[5] = first / last of the ctl[1] words of out
hipError_t launch_wah_edges(const uint32_t* out, uint32_t* ctl, hipStream_t s);
// nt: threads per block (128, 256 or 512); blocks: persistent grid size
hipError_t launch_conj(const ConjArgs& a, int nt, uint32_t blocks, hipStream_t s);
// A batch whose queries are single terms on one index of column `type`
// (qoff[q + 1] = qoff[q] + 1): the persistent range kernel; scan: 0 no
// term runs the scan plan, 1 some do, 2 all do.  nt: 192 or 256.
hipError_t launch_range(const ConjArgs& a, int type, int scan, int nt, uint32_t blocks, hipStream_t s);
uint32_t scan_blocks(uint32_t groups, uint32_t anb);
// counts[i] = sum of parts[i*stride, i*stride + np[i])
hipError_t launch_sum_parts(const uint32_t* parts, uint32_t stride, const uint32_t* np, uint32_t nq,
                            uint32_t* counts, hipStream_t s);
// *count = sum of parts[0 .. n) (one query, n given on the host)
hipError_t launch_sum_parts_n(const uint32_t* parts, uint32_t n, uint32_t* count, hipStream_t s);
// roff[q*stride + t] = sum of parts[q*stride .. q*stride + t), t < n, per
// query q < nq (the row offsets of each query's tiles)
hipError_t launch_scan_parts(const uint32_t* parts, uint32_t stride, uint32_t n, uint32_t nq, uint32_t* roff,
                             hipStream_t s);
hipError_t launch_scan_excl(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp,
                            hipStream_t s);
uint32_t scan_tmp_words(uint32_t n);

} // namespace fbk
