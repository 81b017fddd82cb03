/*
 * fastbit_hip.h -- C ABI of libfastbit_hip.so, the MI355X (gfx950) engine for
 * FastBit's query hot path: WAH bitvector logic, relic/bin range evaluation
 * and index build.
 *
 * Conventions (mirroring src/iapi.h:17-22 of FastBit 2.0.3):
 *   - every function returns an int status, 0 on success, < 0 on error;
 *     no C++ exception crosses this boundary;  fbh_last_error() describes
 *     the last failure of the calling thread;
 *   - handles are opaque; a handle is not shared between host threads
 *     without an external lock (as ibis::bitvector); calls on different
 *     handles may come from several threads: the library holds one lock
 *     for the whole of every call, so such calls run one after another;
 *   - arrays passed in are host memory unless the name says "dev";
 *   - bitvectors cross the boundary in the reference's serialized form
 *     (src/bitvector.cpp:2283-2313): the code words, then active.val when
 *     active.nbits > 0, then active.nbits.
 */
#ifndef FASTBIT_HIP_H
#define FASTBIT_HIP_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FBH_OK 0
#define FBH_ERR_ARG (-1)
#define FBH_ERR_DEVICE (-2)
#define FBH_ERR_NOMEM (-3)
#define FBH_ERR_INTERNAL (-4)
#define FBH_ERR_SIZE (-5)
#define FBH_ERR_NOTFOUND (-6)

/* Logical operators (src/bitvector.h:98-113). */
#define FBH_AND 0
#define FBH_OR 1
#define FBH_XOR 2
#define FBH_MINUS 3

/* The reference routine a binary operator dispatched to
 * (src/bitvector.cpp:1219-1272, 1438-1471, 1582-1606, 1710-1750). */
#define FBH_PATH_C0 1
#define FBH_PATH_D1 2
#define FBH_PATH_D1SWAP 3
#define FBH_PATH_C1 4
#define FBH_PATH_D2 5
#define FBH_PATH_C2 6
#define FBH_PATH_SHORTCUT 7

/* ---------------------------------------------------------------------- */
/* Library */

/* Select the HIP device and create the library stream.  Optional: every
 * other call initialises device 0 on first use.  (ibis::init,
 * src/ibis.h:378-430, plays this role in the reference.) */
int fbh_init(int device);
const char* fbh_last_error(void);
/* Wait for all work queued by the library. */
int fbh_sync(void);

/* ---------------------------------------------------------------------- */
/* Bitvectors: ibis::bitvector (src/bitvector.h:62-318) held in HBM. */

typedef struct fbh_bitvector fbh_bitvector;

/* ibis::bitvector(const array_t<word_t>&) -- src/bitvector.cpp:54-105.
 * Copies the serialized array to the device. */
int fbh_bv_from_serial(const uint32_t* arr, uint32_t n, fbh_bitvector** out);
/* Empty bitvector (src/bitvector.cpp:37). */
int fbh_bv_new(fbh_bitvector** out);
/* ibis::bitvector::set(val, n) -- src/bitvector.cpp:256-276. */
int fbh_bv_set(fbh_bitvector* bv, int val, uint32_t n);
/* Deep copy (src/bitvector.h:680-684). */
int fbh_bv_copy(const fbh_bitvector* src, fbh_bitvector** out);
void fbh_bv_free(fbh_bitvector* bv);

/* write(array_t&) -- src/bitvector.cpp:2283-2313.  *n receives the
 * serialized length; with arr == NULL only the length is returned. */
int fbh_bv_write(fbh_bitvector* bv, uint32_t* arr, uint32_t cap, uint32_t* n);

/* In-place operators followed by compress():
 *   a &= b; a.compress();   (src/bitvector.cpp:1194-1284, 378-469)
 *   a |= b; a.compress();   (1413-1483)
 *   a ^= b; a.compress();   (1557-1618)
 *   a -= b; a.compress();   (1685-1762)
 * The routine chosen by the reference's dispatch decides the result words;
 * *path (may be NULL) receives it (FBH_PATH_*). */
int fbh_bv_op(fbh_bitvector* a, const fbh_bitvector* b, int op, int* path);
/* The out-of-place operators a & b, a | b, a ^ b, a - b
 * (src/bitvector.cpp:1292-1407, 1488-1551, 1624-1679, 1768-1838): *out is a
 * new bitvector (the caller frees it, as the caller deletes the one
 * operator& returns) holding the words the reference's own dispatch writes,
 * not compressed: literal arrays from the c0 / d2 routines (the d2 test is
 * the strict (|a|w + |b|w) * 31 > nbits), the c1 loops' words when one
 * operand is literal, the c2 loops' words, or a copy of one operand for the
 * all0s / all1s shortcuts.  Operands of different sizes go through a copy
 * of the shorter one extended with zeros and the in-place operator, as the
 * reference does.  *path (may be NULL) receives the routine. */
int fbh_bv_op_new(const fbh_bitvector* a, const fbh_bitvector* b, int op,
                  fbh_bitvector** out, int* path);
/* count(mask) -- src/bitvector.cpp:1088-1124: the ones of a AND mask,
 * without forming the result.  The operands must have the same size
 * (FBH_ERR_SIZE otherwise; the reference does not check it). */
int fbh_bv_count(fbh_bitvector* a, fbh_bitvector* mask, uint32_t* cnt);
/* erase(i, j) -- src/bitvector.cpp:1012-1081: removes bits [i, j) (j is
 * clamped to the size; i >= j does nothing; i past the end is an
 * FBH_ERR_ARG error).  The size is counted first (the reference tests j
 * against nbits as it stands), so the result is the reference's for a
 * vector whose size() was taken. */
int fbh_bv_erase(fbh_bitvector* bv, uint32_t i, uint32_t j);

int fbh_bv_compress(fbh_bitvector* bv);   /* src/bitvector.cpp:378-469 */
int fbh_bv_decompress(fbh_bitvector* bv); /* src/bitvector.cpp:474-512 */
int fbh_bv_flip(fbh_bitvector* bv);       /* src/bitvector.cpp:1128-1174 */
/* cnt() and size() -- src/bitvector.h:647-656, do_cnt at
 * src/bitvector.cpp:583-608 (an all-zero vector of more than one word is
 * collapsed to one 0-fill, as the reference does). */
int fbh_bv_cnt(fbh_bitvector* bv, uint32_t* cnt);
int fbh_bv_size(fbh_bitvector* bv, uint32_t* size);
/* adjustSize(nv, nt) growth -- src/bitvector.cpp:4086-4113. */
int fbh_bv_adjust_size(fbh_bitvector* bv, uint32_t nv, uint32_t nt);
/* Ascending positions of the set bits (the indexSet walk,
 * src/bitvector.cpp:4524-4590).  *n receives the count; rids may be NULL
 * to query it; cap bounds what is written. */
int fbh_bv_rids(fbh_bitvector* bv, uint32_t* rids, uint32_t cap, uint32_t* n);
/* Number of code words in m_vec (m_vec.size()). */
int fbh_bv_nwords(const fbh_bitvector* bv, uint32_t* n);

/* Benchmark helper: runs (copy of a) op= b + compress `reps` times on the
 * device with inputs resident, without host synchronisation between reps,
 * and returns the mean wall time per op (ms, HIP events) and the result's
 * word count. */
int fbh_bench_op(const fbh_bitvector* a, const fbh_bitvector* b, int op, int reps,
                 float* ms_per_op, uint32_t* out_words);
/* As fbh_bench_op, cycling over npairs operand pairs (a[i], b[i]) so that
 * the working set exceeds the last-level cache. */
int fbh_bench_op_rot(fbh_bitvector* const* a, fbh_bitvector* const* b, int npairs, int op,
                     int reps, float* ms_per_op, uint32_t* out_words);
/* cnt() (src/bitvector.h:616-656) timed as fbh_bench_op_rot times an op:
 * the device count of a[r % n] `reps` times without host synchronisation;
 * mean ms per count, and the ones of a[0]. */
int fbh_bench_cnt_rot(fbh_bitvector* const* a, int n, int reps, float* ms_per_cnt, uint32_t* ones);

/* ---------------------------------------------------------------------- */
/* Indexes: ibis::relic (equality index, one bitmap per distinct value) and
 * ibis::bin (binned equality index) with their raw column, held in HBM. */

typedef struct fbh_index fbh_index;

#define FBH_RELIC 1
#define FBH_BIN 2

/* Column element types (a subset of ibis::TYPE_T, src/table.h). */
#define FBH_COL_INT32 0
#define FBH_COL_UINT32 1
#define FBH_COL_UINT16 2
#define FBH_COL_FLOAT 3
#define FBH_COL_DOUBLE 4

/* One side of a range lo OP1 x OP2 hi (ibis::qExpr::COMPARE, src/qExpr.h). */
#define FBH_OP_NONE 0
#define FBH_OP_LT 1
#define FBH_OP_LE 2

/* Evaluation plans.  They yield the same bits, except where the reference's
 * own index and scan paths disagree: part::doScan converts the bounds to an
 * integer column's element type (src/part.cpp:11130-11237), so a
 * fractional bound on an integer column can select other rows under
 * FBH_PLAN_SCAN, exactly as it does in the reference. */
#define FBH_PLAN_AUTO 0        /* the ibis heuristic (column::evaluateRange) */
#define FBH_PLAN_LITERAL 1     /* LDS multiway OR into a literal result      */
#define FBH_PLAN_COMPRESSED 2  /* OR merged per tile, compressed result     */
#define FBH_PLAN_SCAN 3        /* raw column compare (part::doScan)          */

/* relic over the distinct values of a column: index::mapValues +
 * relic::construct (src/index.cpp:2552-2611, src/irelic.cpp:878-925). */
int fbh_index_build_relic(const void* col, int type, uint32_t nrows, fbh_index** out);
/* relic(col, card, ind) (src/irelic.cpp:160-211): values 0..card-1. */
int fbh_index_build_relic_card(const uint32_t* ind, uint32_t nrows, uint32_t card,
                               fbh_index** out);
/* Two-level indexes over a relic (src/irelic.h:540-712): the relic is the
 * fine level, and a coarse level groups its bitmaps into ncoarse bins of
 * nearly equal bytes.  ibis::fuzz (interval-equality) keeps
 * ncoarse - ceil(ncoarse/2) + 1 coarse bitmaps, bitmap i holding coarse
 * bins [i, i + ceil(ncoarse/2)); ibis::zona (equality-equality) keeps one
 * per coarse bin.  A range query then reads one or two coarse bitmaps and
 * the fine bitmaps at its edges where the reference's cost rule says that
 * is cheaper (fuzz::evaluate, src/ixfuzz.cpp:872-1095; zona::evaluate,
 * src/ixzona.cpp:713-1017); its rows are those of the relic.  The coarse
 * bitmaps are the canonical WAH of their rows (the reference compresses
 * the result of its own op sequence, which may keep a lone count-1 fill
 * where the canonical form has a literal word). */
#define FBH_ENC_FUZZ 1
#define FBH_ENC_ZONA 2
/* ibis::fuzz(col) (src/ixfuzz.cpp:20-44, coarsen 161-263): ncoarse is the
 * index spec's "ncoarse=" value; 0 (or below 5) picks the reference's
 * default count.  With fewer than 32 distinct values, or where the rule
 * gives fewer than 5 bins, there is no coarse level and queries run on the
 * relic, as in the reference. */
int fbh_index_build_fuzz(const void* col, int type, uint32_t nrows, uint32_t ncoarse, fbh_index** out);
/* ibis::zona(col) (src/ixzona.cpp, coarsen 124-182): 11 coarse bins. */
int fbh_index_build_zona(const void* col, int type, uint32_t nrows, fbh_index** out);
/* The coarse level: the encoding (0 for a plain relic or bin index), the
 * number of coarse bins, their boundaries cbounds[0..ncoarse] (up to cap),
 * the coarse bitmaps' byte offsets as coffset64 holds them (up to ccap;
 * bitvector::bytes() sums for a built index, file positions for one read
 * from a file) and the number of coarse bitmaps. */
int fbh_index_coarse(const fbh_index* ix, int* enc, uint32_t* ncoarse, uint32_t* cbounds, uint32_t cap,
                     int64_t* coffsets, uint32_t ccap, uint32_t* ncbits);
/* Coarse bitmap i as a new bitvector. */
int fbh_index_coarse_bitmap(const fbh_index* ix, uint32_t i, fbh_bitvector** out);
/* What fuzz::evaluate / zona::evaluate does for a range: *option 1 (fine
 * level only) or 2..5 (the coarse options, numbered as in the reference),
 * or one of FBH_TL_*; *coarse_read bit i for every coarse bitmap it reads
 * (the bitmaps a file-backed reference index loads). */
#define FBH_TL_EMPTY 0
#define FBH_TL_FINE 1
#define FBH_TL_ALL 6
#define FBH_TL_EQUAL 7
#define FBH_TL_SAME_BIN 8
int fbh_index_two_level_option(const fbh_index* ix, double lo, int lop, double hi, int rop, int* option,
                               uint64_t* coarse_read);
/* bin with the given boundaries: bin::binning (src/ibin.cpp:3350-3449),
 * empty inner bins dropped.  The bounds are the reference's bounds array
 * (bounds[nobs-1] is DBL_MAX). */
int fbh_index_build_bin(const void* col, int type, uint32_t nrows, const double* bounds,
                        uint32_t nbounds, fbh_index** out);
/* bin with equal-weight boundaries: the "<binning nbins=N equal-weight/>"
 * specification of bin::setBoundaries / scanAndPartition
 * (src/ibin.cpp:2975-3332, 2389-2582): histogram of the distinct values on
 * the device, index::divideCounts (src/index.cpp:3648-4206) and
 * util::compactValue (src/util.cpp:1227-1312) on the host. */
int fbh_index_build_bin_nbins(const void* col, int type, uint32_t nrows, uint32_t nbins,
                              fbh_index** out);
/* The bin boundaries (bounds[nobs-1] == DBL_MAX); *n receives the count. */
int fbh_index_bounds(const fbh_index* ix, double* bounds, uint32_t cap, uint32_t* n);
/* Host-only: the equal-weight boundaries of a histogram (sorted distinct
 * values and their row counts); is_float selects the floating-point rules,
 * col_upper is the column's known upper bound (-DBL_MAX when unknown). */
int fbh_bin_equal_weight(const double* vals, const uint32_t* counts, uint32_t ncnt, uint32_t nbins,
                         int is_float, double col_upper, double* bounds, uint32_t cap,
                         uint32_t* nbounds);
void fbh_index_free(fbh_index* ix);
int fbh_index_info(const fbh_index* ix, uint32_t* nobs, uint32_t* nrows, uint64_t* nwords);
/* fastbit_iapi_deconstruct_index (src/iapi.cpp:1915-1953): keys (relic:
 * values; bin: minval ++ maxval), offsets in words, serialized bitmaps.
 * The arrays are allocated by the library; release them with fbh_free. */
int fbh_index_deconstruct(const fbh_index* ix, double** keys, uint64_t* nkeys,
                          int64_t** offsets, uint64_t* noffsets, uint32_t** bms,
                          uint64_t* nbms);
void fbh_free(void* p);
/* fastbit_iapi_attach_full_index (src/iapi.cpp:2037-2055) over the column
 * it indexes (the array fastbit_iapi_register_array names): the index in
 * the deconstructed form above, bin when nkeys == 2*(noffsets-1) and relic
 * when nkeys+1 == noffsets (column::attachIndex, src/column.cpp:5637-5677;
 * bin boundaries from the minima and maxima as src/ibin.cpp:358-373 makes
 * them).  Unlike the reference, which keeps pointers to the caller's
 * arrays, the words are checked and copied into device memory, so the
 * arrays may be released when this returns.  A bitmap that does not cover
 * exactly nrows rows is an FBH_ERR_ARG error. */
int fbh_index_attach(const void* col, int type, uint32_t nrows, const double* keys, uint64_t nkeys,
                     const int64_t* offsets, uint64_t noffsets, const uint32_t* bms, uint64_t nbms,
                     fbh_index** out);
/* relic::read / bin::read of an index file (src/irelic.cpp:690-786,
 * src/ibin.cpp:468-553): the "#IBIS" header, 4- or 8-byte bitmap
 * positions, the keys, and the serialized bitmaps, mapped from the file and
 * copied to device memory once.  Relic (type 7), bin (type 0), fuzz (type
 * 24; fuzz::read, src/ixfuzz.cpp:1333-1584) and zona (type 25) files; the
 * two-level files carry their coarse level after the relic part.  nrows
 * must equal the row count in the file. */
int fbh_index_read_file(const char* path, const void* col, int type, uint32_t nrows, fbh_index** out);
/* fastbit_iapi_get_number_of_hits (src/iapi.cpp:2011-2017): the ones of
 * index::sumBins(ib, ie) (src/index.cpp:6450-6491) over the index's
 * bitmaps. */
int fbh_index_sum_bins(fbh_index* ix, uint32_t ib, uint32_t ie, uint64_t* hits);
/* fastbit_iapi_resolve_range (src/iapi.cpp:1981-2009): relic::locate /
 * bin::locate of the range. */
int fbh_index_resolve_range(const fbh_index* ix, double lo, int lop, double hi, int rop,
                            uint32_t* cand0, uint32_t* hit0, uint32_t* hit1, uint32_t* cand1);
/* estimateCost (src/ibin.cpp:6531-6569, src/irelic.cpp:1858-1902), the
 * scan cost of column::evaluateRange (src/column.cpp:6102-6111) and the
 * plan the ibis heuristic picks. */
int fbh_index_estimate_cost(const fbh_index* ix, double lo, int lop, double hi, int rop,
                            double* icost, double* scost, int* plan);
/* relic::evaluate / bin::evaluate (src/irelic.cpp:1827-1839,
 * src/ibin.cpp:6614-6713): number of hits and, when result != NULL, the
 * hits as a new (canonical) bitvector the caller frees. */
int fbh_index_evaluate(fbh_index* ix, double lo, int lop, double hi, int rop, int plan,
                       uint64_t* hits, fbh_bitvector** result, int* plan_used);
/* Hit row ids in ascending order (countQuery::getHitRows,
 * src/countQuery.cpp:387-418). */
int fbh_index_evaluate_rids(fbh_index* ix, double lo, int lop, double hi, int rop, int plan,
                            uint32_t* rids, uint64_t cap, uint64_t* n);
/* index::estimate(expr, lower, upper) (src/index.h:191-194): rows of *lower
 * are hits, rows of *upper candidates; an *upper of size 0 (the reference's
 * cleared bitvector) stands for *lower itself.  relic: lower is evaluate's
 * result and upper is cleared (src/irelic.h:58-62).  bin
 * (src/ibin.cpp:6722-6829): lower = the hit bins, upper = lower and the
 * edge bins; no hit bin: lower empty, upper the candidate bins; a range
 * whose bitmap bytes plus edge-bin values cost more than 0.75 bytes per row
 * gives lower = none, upper = every row.  The cost reads the bitmap sizes as
 * the reference's offset64 holds them once sumBins has run (serialized
 * sizes for an index that came with offsets, else bitvector::bytes()).
 * Both bitvectors are new (canonical words); the caller frees them. */
int fbh_index_estimate(fbh_index* ix, double lo, int lop, double hi, int rop, fbh_bitvector** lower,
                       fbh_bitvector** upper);
/* No counterpart in the reference: fbh_index_evaluate for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc,
 * hipFree or other allocating routine.  The literal result (groups + 1
 * words: one word per 31-row group, then the active rows left-aligned in
 * the last word) goes to the caller's device buffer dev_out of out_words
 * words; *hits receives the count.  Plans FBH_PLAN_LITERAL and
 * FBH_PLAN_SCAN; FBH_PLAN_AUTO runs the heuristic's plan, the literal plan
 * where that is the compressed one (the same bits); FBH_PLAN_COMPRESSED is
 * an FBH_ERR_ARG error.  The per-block partial counts live in a static
 * device array of 65,537 words that each GPU holds until the process ends,
 * so indexes of more than 65,536 tiles (2^27 groups) are an FBH_ERR_SIZE
 * error.  Every call on one GPU shares that array: calls that overlap on a
 * GPU (from other processes' threads or captured graphs replayed
 * concurrently) race. */
int fbh_index_evaluate_noalloc(fbh_index* ix, double lo, int lop, double hi, int rop, int plan,
                               uint32_t* dev_out, uint64_t out_words, uint64_t* hits, int* plan_used);
int fbh_index_column_bytes(const fbh_index* ix, uint64_t* bytes);
/* The cost rule of the plan choice (FBH_PLAN_AUTO):
 *   FBH_COST_IAPI       arrays registered through the in-memory API: the
 *                       index when its estimated bytes are below
 *                       nrows/2 + 999 (bord::column::evaluateRange,
 *                       src/bord.cpp:5903-5912); the default;
 *   FBH_COST_PARTITION  data partitions with index files: the index when
 *                       its estimated bytes are below the scan's page bytes
 *                       (column::evaluateRange, src/column.cpp:6102-6111). */
#define FBH_COST_IAPI 0
#define FBH_COST_PARTITION 1
int fbh_index_set_cost_model(fbh_index* ix, int model);

/* Many range queries on one index, queued on the device with one
 * synchronisation at the end; consecutive queries rotate over three
 * streams so that their launches overlap.  plans[i] (or `plan` when plans
 * is NULL)
 * picks the plan per query; hits[i] receives the counts, plans_used[i] the
 * plan that ran, *gpu_ms the device time from the first to the last query.
 * (No counterpart in the reference, which evaluates one selection at a
 * time through fastbit_selection_evaluate, src/iapi.cpp:1383-1418.) */
int fbh_index_evaluate_batch(fbh_index* ix, uint32_t nq, const double* lo, const int* lop,
                             const double* hi, const int* rop, const int* plans, int plan,
                             uint64_t* hits, int* plans_used, float* gpu_ms);
/* Algorithmic bytes a plan moves for one query (bitmap words read, raw
 * values read for the edge check or the scan, literal result written). */
int fbh_index_query_bytes(const fbh_index* ix, double lo, int lop, double hi, int rop, int plan,
                          double* bytes);

/* ---------------------------------------------------------------------- */
/* Conjunctions: countQuery::evaluate of t[0] AND t[1] AND ... over columns of
 * the same rows (src/countQuery.cpp:292-373, 929-1197).  Each term picks its
 * plan as a single range does; terms are evaluated cheapest first, every
 * tile stopping once it is empty. */

typedef struct fbh_term {
    const fbh_index* ix;   /* index (and column) of the term */
    double lo;
    int lop;               /* FBH_OP_* */
    double hi;
    int rop;
} fbh_term;

typedef struct fbh_conj fbh_conj;

/* Prepares a batch: query q ANDs terms[qoff[q] .. qoff[q+1]).  The term
 * descriptors go to the device once and the launch sequence (one
 * persistent kernel over (query, tile) work items and one reduction) is
 * captured in a hipGraph.  (No counterpart in the reference, which
 * evaluates one query at a time; a batch API for many queries.) */
int fbh_conj_prepare(const fbh_term* terms, const uint32_t* qoff, uint32_t nq, int plan,
                     fbh_conj** out);
/* Replays the batch: hits[q] (may be NULL) receives the counts, *gpu_ms
 * (may be NULL) the device time of the replay. */
int fbh_conj_run(fbh_conj* c, uint64_t* hits, float* gpu_ms);
/* Autotuner: times the batch over block sizes, tiles per work item and
 * grid sizes, keeps the fastest and returns it. */
int fbh_conj_tune(fbh_conj* c, int* threads, uint32_t* blocks, uint32_t* tiles_per_item, float* best_ms);
/* The plan each term runs (FBH_PLAN_LITERAL or FBH_PLAN_SCAN), in the
 * caller's term order. */
int fbh_conj_plans(const fbh_conj* c, int* plans);
/* Row ids of every query of the batch (countQuery::getHitRows,
 * src/countQuery.cpp:387-418, for each query): replays the batch for the
 * counts, then one more persistent launch writes every tile's ids.
 * offsets[0 .. nq] (required) receives the exclusive prefix of the counts;
 * the ids of query q, ascending, go to rids[offsets[q] .. offsets[q+1]).
 * At most `cap` ids are written (rids may be NULL: offsets only). */
int fbh_conj_rids(fbh_conj* c, uint64_t* offsets, uint32_t* rids, uint64_t cap);
/* The hit vector of query q (countQuery::getHitVector,
 * src/countQuery.h:64) as a new bitvector in canonical compressed form
 * (the caller frees it). */
int fbh_conj_result(fbh_conj* c, uint32_t q, fbh_bitvector** out);
void fbh_conj_free(fbh_conj* c);
/* A batch of single range queries on one index as a conjunction batch of
 * one-term queries (run, tune and free it with fbh_conj_*): one persistent
 * launch of the range kernel over (query, tiles) work items, captured in a
 * hipGraph and autotuned by fbh_conj_tune; counts only.  plans[i] (or
 * `plan` when plans is NULL) picks each query's plan; FBH_PLAN_AUTO applies
 * the heuristic, and the compressed plan runs as the literal one (the same
 * bits).  On a fuzz / zona index each query is the program of its option
 * (coarse and fine bitmaps combined) in one launch of the expression
 * kernel, and fbh_conj_plans gives one plan per query.  (No counterpart in
 * the reference, which evaluates one selection at a time,
 * src/iapi.cpp:1383-1418.) */
int fbh_index_batch_prepare(fbh_index* ix, uint32_t nq, const double* lo, const int* lop, const double* hi,
                            const int* rop, const int* plans, int plan, fbh_conj** out);
/* One conjunction: *hits receives the count. */
int fbh_conj_evaluate(const fbh_term* terms, uint32_t nterms, int plan, uint64_t* hits);

/* Boolean expressions over range terms: countQuery::doEvaluate over the
 * LOGICAL_AND / OR / XOR / NOT / MINUS tree (src/countQuery.cpp:944-1003)
 * that fastbit_selection_combine builds (src/iapi.cpp:1285-1339; NAND and
 * NOR are NOT over AND and OR there).  Query q is the postfix program
 * prog[poff[q] .. poff[q+1]) over its terms terms[toff[q] .. toff[q+1]):
 * an item i >= 0 pushes the rows of term toff[q] + i, an FBH_EXPR_* item
 * pops two results (NOT: one) and pushes the combination.  At most
 * FBH_EXPR_DEPTH results are on the stack; a program must leave one.
 * NOT complements within the column's rows (every row is valid for arrays
 * registered in memory).  The returned batch runs, tunes and frees like a
 * conjunction batch. */
#define FBH_EXPR_DEPTH 4
#define FBH_EXPR_AND (-1)
#define FBH_EXPR_OR (-2)
#define FBH_EXPR_XOR (-3)
#define FBH_EXPR_NAND (-4)
#define FBH_EXPR_NOR (-5)
#define FBH_EXPR_NOT (-6)
#define FBH_EXPR_MINUS (-7)
int fbh_expr_prepare(const fbh_term* terms, const uint32_t* toff, const int32_t* prog,
                     const uint32_t* poff, uint32_t nq, int plan, fbh_conj** out);

/* ---------------------------------------------------------------------- */
/* Learned plan choice.  The reference picks a plan with fixed cost rules
 * (column::evaluateRange, src/column.cpp:6102-6111; bin::sumBins,
 * src/index.cpp:6640-6783); this is a learned stand-in for that choice,
 * with no counterpart in the reference.  Every plan yields the same bits,
 * so the choice affects only the time.
 *
 * Features of a range query (FBH_NFEATURES floats, from index metadata):
 *   0 log2(1 + words of the bitmaps in the hit range)
 *   1 log2(1 + words of the bitmaps outside it)
 *   2 log2(1 + words the literal plan reads, edge bitmaps included)
 *   3 bin span: hit bitmaps / bitmaps
 *   4 density: rows of the hit bitmaps / rows
 *   5 log2(1 + rows of the edge bitmaps)
 *   6 log2(1 + bitmaps in the OR set)
 *   7 1 when the OR set is the complement of the hit range
 *   8 log2(1 + rows)   9 log2(1 + bitmaps)
 *  10 1 for a floating-point column   11 element size in bytes */
#define FBH_NFEATURES 12
#define FBH_NPLANS 3
int fbh_index_plan_features(const fbh_index* ix, uint32_t nq, const double* lo, const int* lop,
                            const double* hi, const int* rop, float* feats);
/* Times every query under the three plans, back to back on the device:
 * ms[3*i + plan-1] is the device time (HIP events around the query's
 * launches) of query i under the literal, compressed and scan plans, and
 * hits[3*i + plan-1] its count.  The literal and scan plans write the
 * literal result, the compressed plan the compressed one.  A negative time
 * would mark a lower bound; the device plans never produce one.  cap is
 * kept for the interface and not used. */
int fbh_index_time_plans(fbh_index* ix, uint32_t nq, const double* lo, const int* lop, const double* hi,
                         const int* rop, float cap, float* ms, uint64_t* hits);

/* The cost model: an FP32 MLP 12 -> 64 -> 64 -> 3 (ReLU) whose outputs are
 * the predicted log2 device times (microseconds) of the literal,
 * compressed and scan plans; trained and run by the library's own kernels.
 * Parameters, flat: W1[64][12] b1[64] W2[64][64] b2[64] W3[3][64] b3[3]. */
typedef struct fbh_planner fbh_planner;
#define FBH_PLANNER_NPARAM 5187
/* He-uniform weights from seed, zero biases, identity normalisation. */
int fbh_planner_new(uint64_t seed, fbh_planner** out);
void fbh_planner_free(fbh_planner* p);
int fbh_planner_set_params(fbh_planner* p, const float* params, uint32_t n);
int fbh_planner_get_params(const fbh_planner* p, float* params, uint32_t n);
/* Input normalisation x' = (x - mu) / sd (sd <= 0 counts as 1). */
int fbh_planner_set_norm(fbh_planner* p, const float* mu, const float* sd);
int fbh_planner_get_norm(const fbh_planner* p, float* mu, float* sd);
/* Scores of n samples (x: n x 12, out: n x 3). */
int fbh_planner_forward(fbh_planner* p, const float* x, uint32_t n, float* out);
/* Loss and gradient over n samples at the current parameters:
 * L = sum_i sum_o w_io e_io^2 / sum |mask|, e = score - y, where mask > 0
 * is an exact label of weight mask, mask < 0 a lower bound of weight -mask
 * (e counts only when negative), 0 no label. */
int fbh_planner_grad(fbh_planner* p, const float* x, const float* y, const float* mask, uint32_t n,
                     float* loss, float* grad);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over shuffled mini-batches of the
 * given size; normalisation and the in-range box of the features are set
 * from x first.  epoch_loss[e] (may be NULL) receives each epoch's mean
 * loss. */
int fbh_planner_train(fbh_planner* p, const float* x, const float* y, const float* mask, uint32_t n,
                      uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float* epoch_loss);
/* One Adam step on one batch of n samples (the test of the update). */
int fbh_planner_step(fbh_planner* p, const float* x, const float* y, const float* mask, uint32_t n, float lr);
/* Plans of n feature rows: the lowest score, or fallback[i] when a
 * feature lies outside the training box (and for every row of an
 * untrained planner).  *nfallback (may be NULL) counts the fallbacks. */
int fbh_planner_choose(fbh_planner* p, const float* x, uint32_t n, const int* fallback, int* plans,
                       uint32_t* nfallback);
/* Features, the ibis heuristic's plan as the fallback and the learned
 * choice for queries on one index. */
int fbh_index_choose_plans(fbh_index* ix, fbh_planner* p, uint32_t nq, const double* lo, const int* lop,
                           const double* hi, const int* rop, int* plans, uint32_t* nfallback);

#ifdef __cplusplus
}
#endif
#endif /* FASTBIT_HIP_H */
