// C ABI of the relic / bin part of libfastbit_hip.so (include/fastbit_hip.h).
//
// Build: keys on the GPU (relic: position of the value among the sorted
// distinct values; bin: bin::locate), a stable radix sort of (key, row), one
// canonical WAH bitmap per key, concatenated in the reference's in-memory
// serialized layout; then the tile offset table used by range evaluation.
//
// Evaluate (relic::evaluate, src/irelic.cpp:1827-1839; bin::evaluate,
// src/ibin.cpp:6614-6713): locate on the host, then one of three plans that
// all yield the same bits -- LITERAL (LDS multiway OR + edge check),
// COMPRESSED (pairwise WAH ORs of the bitmaps), SCAN (raw column compare).
#include "../../include/fastbit_hip.h"
#include "runtime.h"
#include "launch.h"
#include "bitvector_impl.h"
#include "index_dev.h"
#include "index_host.h"
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

using namespace fbh_rt;

namespace {

constexpr uint32_t MAXBITS = 31;
inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
int fail(int code, const std::string& msg) { set_error(msg); return code; }

size_t col_esize(int type) {
    switch (type) {
    case FBH_COL_INT32: case FBH_COL_UINT32: case FBH_COL_FLOAT: return 4;
    case FBH_COL_UINT16: return 2;
    case FBH_COL_DOUBLE: return 8;
    default: return 0;
    }
}

int dev_type(int type) {
    switch (type) {
    case FBH_COL_INT32: return fbk::COL_I32;
    case FBH_COL_UINT32: return fbk::COL_U32;
    case FBH_COL_UINT16: return fbk::COL_U16;
    case FBH_COL_FLOAT: return fbk::COL_F32;
    default: return fbk::COL_F64;
    }
}

bool is_int_type(int type) {
    return type == FBH_COL_INT32 || type == FBH_COL_UINT32 || type == FBH_COL_UINT16;
}

// RAII device buffer from the library allocator.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    int alloc(size_t cnt) {
        n = cnt;
        p = (T*)dev_alloc(sizeof(T) * (cnt ? cnt : 1));
        return p ? FBH_OK : fail(FBH_ERR_NOMEM, "device allocation failed");
    }
    ~DBuf() { dev_free(p); }
    T* release() { T* q = p; p = nullptr; return q; }
};

int stream_sync() { return check(hipStreamSynchronize(stream()), "sync"); }

template <class T>
int to_dev(T* d, const T* h, size_t n) {
    if (!n) return FBH_OK;
    return check(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, stream()), "upload");
}

template <class T>
int to_host(T* h, const T* d, size_t n) {
    if (!n) return FBH_OK;
    int rc = check(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, stream()), "download");
    return rc ? rc : stream_sync();
}

} // namespace

struct fbh_index {
    int kind = 0;                 // FBH_RELIC / FBH_BIN
    int type = 0;                 // column type
    uint32_t nrows = 0, groups = 0, anb = 0, nobs = 0;
    std::vector<double> keys;     // relic: distinct values; bin: bounds
    std::vector<double> minval, maxval;
    std::vector<uint64_t> off;    // serialized offsets (words), nobs+1
    std::vector<uint32_t> nw;     // code words per bitmap (0 and no trailer: null)
    uint32_t* words = nullptr;    // serialized bitmaps (device)
    uint64_t nwords = 0;
    uint64_t* doff = nullptr;
    uint32_t* dnw = nullptr;
    uint32_t ntiles = 0;
    uint32_t* tw = nullptr;       // tiled copy of the bitmaps (literal plan)
    uint32_t* TW = nullptr;       // segment offsets, ntiles*nobs + 1
    uint64_t ntw = 0;
    void* col = nullptr;          // raw column (device)
    double* dbounds = nullptr;    // bin bounds (device)
    std::vector<double> cumbytes; // prefix sums of bytes(b)
    std::vector<double> cumser;   // prefix sums of serialized bytes
    std::vector<double> cumrows;  // prefix sums of rows per bitmap
    bool offsets_built = false;   // the reference constructor fills offset64
    int cost_model = FBH_COST_IAPI;
    double cmin = DBL_MAX, cmax = -DBL_MAX;   // column min / max
    // two-level encodings over a relic (ibis::fuzz / ibis::zona): the coarse
    // bins, the byte offsets the cost rules read, and the coarse bitmaps as
    // an index of their own (its words, tiled copy and tables)
    int enc = 0;                          // FBH_ENC_FUZZ, FBH_ENC_ZONA or 0
    std::vector<uint32_t> cbounds;        // coarse bin c: fine bitmaps [cbounds[c], cbounds[c+1])
    std::vector<int64_t> toff;            // offset64 of the fine level (bytes)
    std::vector<int64_t> tcoff;           // coffset64 of the coarse bitmaps (bytes)
    fbh_index* coarse = nullptr;
    ~fbh_index() {
        dev_free(words); dev_free(doff); dev_free(dnw); dev_free(tw); dev_free(TW); dev_free(col);
        dev_free(dbounds);
        delete coarse;
    }
    bool two_level() const { return coarse != nullptr && !cbounds.empty(); }
    bool null_bitmap(uint32_t b) const { return off[b + 1] == off[b]; }
    // ibis::bitvector::bytes(): code words * 4 + sizeof(bitvector) (40 on x86-64)
    double bytes(uint32_t b) const { return null_bitmap(b) ? 0.0 : 4.0 * nw[b] + 40.0; }
};

namespace {

// Stable sort of (key, row) by key; `bits` key bits are sorted.
int sort_pairs(uint32_t* keys, uint32_t n, uint32_t bits, DBuf<uint32_t>& okeys,
               DBuf<uint32_t>& orows) {
    int rc = okeys.alloc(n);
    if (!rc) rc = orows.alloc(n);
    if (rc) return rc;
    DBuf<uint32_t> k2, r2, tmp;
    rc = k2.alloc(n);
    if (!rc) rc = r2.alloc(n);
    const uint32_t nb = fbk::rs_blocks(n);
    if (!rc) rc = tmp.alloc(512ull * nb + fbk::scan_tmp_words(256u * nb) + 16);
    if (rc) return rc;
    const uint32_t passes = bits == 0 ? 1 : cdiv(bits, 8);
    const uint32_t* ki = keys;
    const uint32_t* vi = nullptr;
    // alternate buffers so that the last pass lands in okeys/orows
    uint32_t* bufk[2] = {okeys.p, k2.p};
    uint32_t* bufv[2] = {orows.p, r2.p};
    int cur = (passes % 2 == 1) ? 0 : 1;
    for (uint32_t p = 0; p < passes; ++p) {
        rc = check(fbk::launch_rs_pass(ki, vi, n, 8 * p, bufk[cur], bufv[cur], tmp.p, stream()),
                   "radix pass");
        if (rc) return rc;
        ki = bufk[cur]; vi = bufv[cur];
        cur ^= 1;
    }
    return FBH_OK;
}

uint32_t bits_for(uint32_t maxkey) {
    uint32_t b = 0;
    while (b < 32 && (maxkey >> b) != 0) ++b;
    return b;
}

// Encode bitmaps from sorted rows; fills idx words/offsets/nw (+min/max).
// Two passes of the chunk-parallel encoder: the first counts every chunk's
// code words (and ORs the active bits, reduces the value ranges), the host
// lays the bitmaps out (dropping empty inner bins, deciding which bitmaps
// are stored decompressed), the second writes the words in place.
int encode_bitmaps(fbh_index* ix, const uint32_t* skeys_d, const uint32_t* srows_d, uint32_t nsorted,
                   bool with_minmax, bool drop_empty_bins, bool unpack_heavy) {
    const uint32_t nbm = ix->nobs;
    // bitmap starts in the sorted key order (rows without a bitmap sort last)
    std::vector<uint32_t> start(nbm + 1);
    DBuf<uint32_t> dstart;
    int rc = dstart.alloc(nbm + 1);
    if (!rc) rc = check(fbk::launch_starts(skeys_d, nsorted, nbm, dstart.p, stream()), "starts");
    if (!rc) rc = to_host(start.data(), dstart.p, nbm + 1);
    if (rc) return rc;
    // chunks of at most ENC_ROWS rows, one at least per bitmap
    std::vector<uint32_t> cbm, cstart, cfirst(nbm + 1);
    for (uint32_t b = 0; b < nbm; ++b) {
        cfirst[b] = (uint32_t)cbm.size();
        const uint32_t n = start[b + 1] - start[b];
        const uint32_t k = n ? (n + fbk::ENC_ROWS - 1) / fbk::ENC_ROWS : 1u;
        for (uint32_t j = 0; j < k; ++j) { cbm.push_back(b); cstart.push_back(start[b] + j * fbk::ENC_ROWS); }
    }
    const uint32_t nchunk = (uint32_t)cbm.size();
    cfirst[nbm] = nchunk;
    DBuf<uint32_t> dcbm, dcstart, dccnt, daval;
    DBuf<double> dmin, dmax;
    if ((rc = dcbm.alloc(nchunk)) || (rc = dcstart.alloc(nchunk)) || (rc = dccnt.alloc(nchunk)) ||
        (rc = daval.alloc(nbm ? nbm : 1)))
        return rc;
    if (with_minmax && ((rc = dmin.alloc(nchunk)) || (rc = dmax.alloc(nchunk)))) return rc;
    if ((rc = to_dev(dcbm.p, cbm.data(), nchunk)) || (rc = to_dev(dcstart.p, cstart.data(), nchunk))) return rc;
    if (nbm && (rc = check(fbk::launch_zero_u32(daval.p, nbm, stream()), "zero"))) return rc;
    fbk::EncArgs ea;
    memset(&ea, 0, sizeof(ea));
    ea.rows = srows_d; ea.start = dstart.p; ea.cbm = dcbm.p; ea.cstart = dcstart.p; ea.nchunk = nchunk;
    ea.groups = ix->groups; ea.anb = ix->anb;
    ea.ccnt = dccnt.p; ea.aval = daval.p;
    ea.cmin = with_minmax ? dmin.p : nullptr; ea.cmax = with_minmax ? dmax.p : nullptr;
    rc = check(fbk::launch_enc_count(ea, dev_type(ix->type), with_minmax ? ix->col : nullptr, stream()),
               "encode count");
    std::vector<uint32_t> ccnt(nchunk);
    std::vector<double> cmn, cmx;
    if (!rc) rc = to_host(ccnt.data(), dccnt.p, nchunk);
    if (!rc && with_minmax) {
        cmn.resize(nchunk); cmx.resize(nchunk);
        if (!(rc = to_host(cmn.data(), dmin.p, nchunk))) rc = to_host(cmx.data(), dmax.p, nchunk);
    }
    if (rc) return rc;
    std::vector<uint32_t> nw(nbm, 0);
    std::vector<double> mn(nbm, DBL_MAX), mx(nbm, -DBL_MAX);
    for (uint32_t c = 0; c < nchunk; ++c) {
        const uint32_t b = cbm[c];
        nw[b] += ccnt[c];
        if (with_minmax && start[b + 1] > start[b]) { mn[b] = std::min(mn[b], cmn[c]); mx[b] = std::max(mx[b], cmx[c]); }
    }
    // bins: drop empty ones except the first and the last
    // (src/ibin.cpp:3397-3428); an empty kept bin is a null bitmap
    std::vector<uint32_t> keep;
    std::vector<bool> nullbm;
    if (drop_empty_bins) {
        for (uint32_t b = 0; b < nbm; ++b) {
            const bool empty = start[b + 1] == start[b];
            if (!empty || b == 0 || b + 1 == nbm) { keep.push_back(b); nullbm.push_back(empty); }
        }
    } else {
        for (uint32_t b = 0; b < nbm; ++b) { keep.push_back(b); nullbm.push_back(false); }
    }
    const uint32_t nk = (uint32_t)keep.size();
    std::vector<uint64_t> off(nk + 1);
    std::vector<uint32_t> knw(nk);
    // index::optionalUnpack with no <compressing> directive
    // (src/index.cpp:8803-8854): a bitmap whose bytes() exceed
    // max(nrows/9, sizeof(bitvector)+12) is stored decompressed.
    const uint64_t bar1 = std::max<uint64_t>(ix->nrows / 9u, 40u + 12u);
    const uint32_t trailer = ix->anb > 0 ? 2u : 1u;
    std::vector<uint64_t> wpos(nchunk, ~0ull), boff(nbm, 0);
    std::vector<uint32_t> bnw(nbm, 0);
    std::vector<uint8_t> unpack(nbm, 0);
    bool any_unpack = false;
    off[0] = 0;
    for (uint32_t k = 0; k < nk; ++k) {
        const uint32_t b = keep[k];
        knw[k] = nullbm[k] ? 0u : nw[b];
        if (unpack_heavy && !nullbm[k] && 4ull * nw[b] + 40u > bar1 && nw[b] != ix->groups) {
            unpack[b] = 1;
            knw[k] = ix->groups;
            any_unpack = true;
        }
        off[k + 1] = off[k] + (nullbm[k] ? 0u : knw[k] + trailer);
        if (nullbm[k]) continue;
        boff[b] = off[k];
        bnw[b] = knw[k];
        uint64_t p = off[k];
        for (uint32_t c = cfirst[b]; c < cfirst[b + 1]; ++c) { wpos[c] = p; p += ccnt[c]; }
    }
    DBuf<uint64_t> dwpos, dboff;
    DBuf<uint32_t> dbnw, words;
    DBuf<uint8_t> dunpack;
    if ((rc = dwpos.alloc(nchunk)) || (rc = dboff.alloc(nbm ? nbm : 1)) || (rc = dbnw.alloc(nbm ? nbm : 1)) ||
        (rc = words.alloc(off[nk] + 1)))
        return rc;
    if ((rc = to_dev(dwpos.p, wpos.data(), nchunk)) || (nbm && (rc = to_dev(dboff.p, boff.data(), nbm))) ||
        (nbm && (rc = to_dev(dbnw.p, bnw.data(), nbm))))
        return rc;
    if (any_unpack) {
        if ((rc = dunpack.alloc(nbm)) || (rc = to_dev(dunpack.p, unpack.data(), nbm))) return rc;
        // the decompressed bitmaps get a literal per group: zero them first
        for (uint32_t b = 0; b < nbm; ++b)
            if (unpack[b] && (rc = check(fbk::launch_zero_u32(words.p + boff[b], ix->groups, stream()), "zero")))
                return rc;
    }
    rc = check(fbk::launch_zero_u32(words.p + off[nk], 1, stream()), "zero");
    ea.out = words.p; ea.wpos = dwpos.p; ea.boff = dboff.p; ea.bnw = dbnw.p;
    ea.unpack = any_unpack ? dunpack.p : nullptr;
    if (!rc) rc = check(fbk::launch_enc_write(ea, stream()), "encode write");
    if (!rc) rc = stream_sync();
    if (rc) return rc;
    // install
    ix->cumrows.assign(nk + 1, 0.0);
    for (uint32_t k = 0; k < nk; ++k)
        ix->cumrows[k + 1] = ix->cumrows[k] + (double)(start[keep[k] + 1] - start[keep[k]]);
    ix->nobs = nk;
    ix->off = off;
    ix->nw = knw;
    ix->nwords = off[nk];
    ix->words = words.release();
    if (ix->kind == FBH_BIN) {
        std::vector<double> nb(nk), nmin(nk), nmax(nk);
        for (uint32_t k = 0; k < nk; ++k) {
            nb[k] = ix->keys[keep[k]];
            nmin[k] = mn[keep[k]];
            nmax[k] = mx[keep[k]];
        }
        ix->keys.swap(nb);
        ix->minval.swap(nmin);
        ix->maxval.swap(nmax);
    } else {
        std::vector<double> nv(nk);
        for (uint32_t k = 0; k < nk; ++k) nv[k] = ix->keys[keep[k]];
        ix->keys.swap(nv);
    }
    return FBH_OK;
}

// Device copies of offsets / word counts and the tile offset table.
int finish_index(fbh_index* ix) {
    int rc;
    ix->doff = (uint64_t*)dev_alloc(8ull * (ix->nobs + 1));
    ix->dnw = (uint32_t*)dev_alloc(4ull * (ix->nobs + 1));
    if (!ix->doff || !ix->dnw) return fail(FBH_ERR_NOMEM, "device allocation failed");
    if ((rc = to_dev(ix->doff, ix->off.data(), ix->nobs + 1)) || (rc = to_dev(ix->dnw, ix->nw.data(), ix->nobs)))
        return rc;
    ix->cumbytes.assign(ix->nobs + 1, 0.0);
    ix->cumser.assign(ix->nobs + 1, 0.0);
    for (uint32_t b = 0; b < ix->nobs; ++b) {
        ix->cumbytes[b + 1] = ix->cumbytes[b] + ix->bytes(b);
        ix->cumser[b + 1] = ix->cumser[b] + 4.0 * (double)(ix->off[b + 1] - ix->off[b]);
    }
    if (ix->kind == FBH_BIN && ix->nobs) {
        ix->dbounds = (double*)dev_alloc(8ull * ix->nobs);
        if (!ix->dbounds) return fail(FBH_ERR_NOMEM, "device allocation failed");
        if ((rc = to_dev(ix->dbounds, ix->keys.data(), ix->nobs))) return rc;
    }
    // the tile offset table holds absolute 32-bit word indices
    if (ix->nwords >= 0xFFFFFFFFull) return fail(FBH_ERR_SIZE, "index larger than 2^32 words");
    ix->ntiles = ix->groups / (uint32_t)fbk::QT + 1;
    // tile offset table, then the tiled copy of the bitmaps built from it
    const size_t nseg = (size_t)ix->ntiles * ix->nobs;
    if (nseg + 1 >= 0xFFFFFFFFull) return fail(FBH_ERR_SIZE, "too many tile segments");
    DBuf<uint2> toff;
    DBuf<uint32_t> len, tmp;
    if ((rc = toff.alloc((size_t)(ix->ntiles + 1) * (ix->nobs ? ix->nobs : 1))) ||
        (rc = len.alloc(nseg + 1)) || (rc = tmp.alloc(fbk::scan_tmp_words((uint32_t)nseg + 1) + 4)))
        return rc;
    ix->TW = (uint32_t*)dev_alloc(4ull * (nseg + 1));
    if (!ix->TW) return fail(FBH_ERR_NOMEM, "device allocation failed");
    {
        // tile offsets over chunks of code words: their groups, a per-bitmap
        // scan here, then the rows
        std::vector<uint32_t> tb, tw0;
        for (uint32_t b = 0; b < ix->nobs; ++b)
            for (uint32_t w = 0; w < ix->nw[b]; w += fbk::TOFF_WORDS) { tb.push_back(b); tw0.push_back(w); }
        const uint32_t nch = (uint32_t)tb.size();
        DBuf<uint32_t> dtb, dtw0, dclen, dcg0, dtot;
        if ((rc = dtb.alloc(nch ? nch : 1)) || (rc = dtw0.alloc(nch ? nch : 1)) || (rc = dclen.alloc(nch ? nch : 1)) ||
            (rc = dcg0.alloc(nch ? nch : 1)) || (rc = dtot.alloc(ix->nobs ? ix->nobs : 1)))
            return rc;
        if (nch && ((rc = to_dev(dtb.p, tb.data(), nch)) || (rc = to_dev(dtw0.p, tw0.data(), nch)))) return rc;
        rc = check(fbk::launch_toff_len(ix->words, ix->doff, ix->dnw, dtb.p, dtw0.p, nch, dclen.p, stream()),
                   "tile offset lengths");
        std::vector<uint32_t> clen(nch), cg0(nch), tot(ix->nobs, 0);
        if (!rc && nch) rc = to_host(clen.data(), dclen.p, nch);
        if (rc) return rc;
        for (uint32_t c = 0; c < nch; ++c) { cg0[c] = tot[tb[c]]; tot[tb[c]] += clen[c]; }
        if (nch && (rc = to_dev(dcg0.p, cg0.data(), nch))) return rc;
        if (ix->nobs && (rc = to_dev(dtot.p, tot.data(), ix->nobs))) return rc;
        rc = check(fbk::launch_toff_rows(ix->words, ix->doff, ix->dnw, dtb.p, dtw0.p, dcg0.p, nch, ix->nobs,
                                         ix->ntiles, (uint32_t)fbk::QT, toff.p, stream()), "tile offsets");
        if (!rc) rc = check(fbk::launch_toff_tail(ix->doff, ix->dnw, dtot.p, ix->nobs, ix->ntiles, (uint32_t)fbk::QT,
                                                  ix->groups, toff.p, stream()), "tile offset tails");
        if (!rc) rc = stream_sync();        // the chunk tables are freed on return
    }
    if (!rc) rc = check(fbk::launch_tseg_len(toff.p, ix->nobs, ix->ntiles, ix->groups, ix->anb, len.p,
                                             stream()), "segment lengths");
    if (!rc) rc = check(fbk::launch_scan_excl(len.p, ix->TW, (uint32_t)nseg + 1, tmp.p, stream()), "segment scan");
    uint32_t total = 0;
    if (!rc) rc = to_host(&total, ix->TW + nseg, 1);
    if (rc) return rc;
    ix->ntw = total;
    // four words of padding: the literal plan reads whole 16-byte groups
    ix->tw = (uint32_t*)dev_alloc(4ull * (total + 4));
    if (!ix->tw) return fail(FBH_ERR_NOMEM, "device allocation failed");
    rc = check(fbk::launch_tseg_fill(ix->words, toff.p, ix->TW, ix->nobs, ix->ntiles, ix->groups, ix->anb,
                                     ix->tw, stream()), "tiled copy");
    if (rc) return rc;
    return stream_sync();
}

int upload_column(fbh_index* ix, const void* col, int type, uint32_t nrows) {
    const size_t es = col_esize(type);
    if (!es) return fail(FBH_ERR_ARG, "unknown column type");
    ix->type = type;
    ix->nrows = nrows;
    ix->groups = nrows / MAXBITS;
    ix->anb = nrows - ix->groups * MAXBITS;
    ix->col = dev_alloc(es * (nrows ? nrows : 1));
    if (!ix->col) return fail(FBH_ERR_NOMEM, "device allocation failed");
    if (nrows) {
        int rc = check(hipMemcpyAsync(ix->col, col, es * nrows, hipMemcpyHostToDevice, stream()),
                       "column upload");
        if (rc) return rc;
    }
    return FBH_OK;
}

// Distinct values of the column (sorted), computed on the GPU: order-
// preserving 32-bit keys, radix sort, host pass over the sorted keys.
// With counts != nullptr also the number of rows of every value (the
// histogram of index::mapValues, src/index.cpp:3411-3479).
int distinct_values(fbh_index* ix, std::vector<double>& vals, std::vector<uint32_t>* counts = nullptr) {
    const uint32_t n = ix->nrows;
    if (ix->type == FBH_COL_DOUBLE) return fail(FBH_ERR_ARG, "relic over double columns is not supported");
    DBuf<uint32_t> k, uq, tmp;
    int rc;
    if ((rc = k.alloc(n)) || (rc = uq.alloc(n)) || (rc = tmp.alloc(2ull * n + fbk::scan_tmp_words(n) + 4)))
        return rc;
    rc = check(fbk::launch_order_keys(ix->col, dev_type(ix->type), n, k.p, stream()), "order keys");
    if (rc) return rc;
    DBuf<uint32_t> sk, sr;
    if ((rc = sort_pairs(k.p, n, 32, sk, sr))) return rc;
    rc = check(fbk::launch_unique(sk.p, n, uq.p, tmp.p, stream()), "unique");
    if (rc) return rc;
    // number of distinct keys = last exclusive position + last flag
    uint32_t h[2] = {0, 0};
    if (n) {
        if ((rc = to_host(&h[0], tmp.p + n + (n - 1), 1)) || (rc = to_host(&h[1], tmp.p + (n - 1), 1)))
            return rc;
    }
    const uint32_t nd = n ? h[0] + h[1] : 0;
    std::vector<uint32_t> s(nd);
    if ((rc = to_host(s.data(), uq.p, nd))) return rc;
    if (counts) {
        // first sorted position of every distinct key -> run lengths
        DBuf<uint32_t> first;
        if ((rc = first.alloc(nd + 1))) return rc;
        rc = check(fbk::launch_unique_first(tmp.p, tmp.p + n, n, first.p, stream()), "unique first");
        std::vector<uint32_t> f(nd + 1);
        if (!rc) rc = to_host(f.data(), first.p, nd);
        if (rc) return rc;
        f[nd] = n;
        counts->resize(nd);
        for (uint32_t i = 0; i < nd; ++i) (*counts)[i] = f[i + 1] - f[i];
    }
    vals.resize(nd);
    for (uint32_t i = 0; i < nd; ++i) {
        const uint32_t u = s[i];
        switch (ix->type) {
        case FBH_COL_INT32: vals[i] = (double)(int32_t)(u ^ 0x80000000u); break;
        case FBH_COL_UINT32: case FBH_COL_UINT16: vals[i] = (double)u; break;
        default: {
            const uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
            float f; memcpy(&f, &b, 4); vals[i] = (double)f; break;
        }
        }
    }
    return FBH_OK;
}

int build_common(fbh_index* ix, const uint32_t* dkeys, uint32_t maxkey, bool minmax, bool drop,
                 bool unpack_heavy) {
    DBuf<uint32_t> sk, sr;
    int rc = sort_pairs(const_cast<uint32_t*>(dkeys), ix->nrows, bits_for(maxkey), sk, sr);
    if (rc) return rc;
    rc = encode_bitmaps(ix, sk.p, sr.p, ix->nrows, minmax, drop, unpack_heavy);
    if (rc) return rc;
    return finish_index(ix);
}

} // namespace

// ---------------------------------------------------------------------------
// Evaluation.
namespace {

struct Locate {
    uint32_t cand0 = 0, hit0 = 0, hit1 = 0, cand1 = 0;
};

// relic::locate / bin::locate as they return (an empty range can come
// back with hit1 < hit0)
void locate_raw(const fbh_index* ix, double lo, int lop, double hi, int rop, Locate& L) {
    if (ix->kind == FBH_RELIC) {
        fbh_ix::relic_locate_range(ix->keys, lo, lop, hi, rop, L.hit0, L.hit1);
        L.cand0 = L.hit0; L.cand1 = L.hit1;
    } else {
        fbh_ix::bin_locate_range(ix->keys, ix->minval, ix->maxval, lo, lop, hi, rop, L.cand0,
                                 L.cand1, L.hit0, L.hit1);
    }
}

void locate(const fbh_index* ix, double lo, int lop, double hi, int rop, Locate& L) {
    locate_raw(ix, lo, lop, hi, rop, L);
    if (L.hit1 < L.hit0) L.hit1 = L.hit0;
}

// sumBins' choice between OR of [ib,ie) and the complement of the OR of
// the outside (src/index.cpp:6450-6491).  The measure is offset64: the
// serialized sizes where a constructor or a file filled it, else the
// bitmap bytes sumBins fills it with.
struct SumPlan {
    int kind;         // 0 none (all zero), 1 all ones, 2 straight, 3 complement
};

SumPlan sum_plan(const fbh_index* ix, uint32_t ib, uint32_t ie) {
    SumPlan p{2};
    if (ie > ix->nobs) ie = ix->nobs;
    if (ib >= ie) { p.kind = 0; return p; }
    const std::vector<double>& m = ix->offsets_built ? ix->cumser : ix->cumbytes;
    const double all = m[ix->nobs];
    const double mid = m[ie] - m[ib];
    if (mid == 0) p.kind = 0;
    else if (all == mid) p.kind = 1;
    else p.kind = (mid <= std::floor(all / 2)) ? 2 : 3;
    return p;
}

void fill_query(const fbh_index* ix, fbk::QueryArgs& q, double lo, int lop, double hi, int rop) {
    memset(&q, 0, sizeof(q));
    q.tw = ix->tw; q.TW = ix->TW; q.ntiles = ix->ntiles;
    q.nobs = ix->nobs; q.groups = ix->groups; q.anb = ix->anb;
    q.cand0 = fbk::NOKEY; q.cand1 = fbk::NOKEY;
    q.col = ix->col;
    q.is_int = is_int_type(ix->type) ? 1 : 0;
    int l = lop, r = rop;
    int64_t ilo = 0, ihi = 0;
    if (q.is_int) {
        if (ix->type == FBH_COL_INT32) fbh_ix::int_bounds<int32_t>(lo, l, hi, r, ilo, ihi);
        else if (ix->type == FBH_COL_UINT32) fbh_ix::int_bounds<uint32_t>(lo, l, hi, r, ilo, ihi);
        else fbh_ix::int_bounds<uint16_t>(lo, l, hi, r, ilo, ihi);
    }
    q.lop = l; q.rop = r; q.lo = lo; q.hi = hi; q.ilo = ilo; q.ihi = ihi;
}

// The bitmap ranges of sumBins(ib, ie) as the literal plan ORs them.
void set_sum_ranges(const fbh_index* ix, fbk::QueryArgs& q, uint32_t ib, uint32_t ie) {
    const SumPlan sp = sum_plan(ix, ib, ie);
    q.r0 = q.r1 = q.r2 = q.r3 = 0;
    if (sp.kind == 1) { q.r0 = 0; q.r1 = 0; q.flip = 1; }
    else if (sp.kind == 2) { q.r0 = ib; q.r1 = std::min(ie, ix->nobs); }
    else if (sp.kind == 3) { q.r0 = 0; q.r1 = ib; q.r2 = std::min(ie, ix->nobs); q.r3 = ix->nobs; q.flip = 1; }
}

// Per-block partial counts of evaluated queries (one row per query), summed
// on the device by one block per query.
struct PartCounts {
    DBuf<uint32_t> parts, np, counts;
    uint32_t stride = 0;
    int alloc(const fbh_index* ix, uint32_t nq) {
        stride = std::max(ix->ntiles, fbk::scan_blocks(ix->groups, ix->anb));
        int rc = parts.alloc((size_t)stride * (nq ? nq : 1));
        if (!rc) rc = np.alloc(nq ? nq : 1);
        if (!rc) rc = counts.alloc(nq ? nq : 1);
        return rc;
    }
    uint32_t* row(uint32_t i) { return parts.p + (size_t)i * stride; }
    // queues the sums (host nparts must stay alive until the stream syncs)
    int launch_sum(const uint32_t* nparts, uint32_t nq) {
        int rc = to_dev(np.p, nparts, nq);
        if (!rc) rc = check(fbk::launch_sum_parts(parts.p, stride, np.p, nq, counts.p, stream()), "sum parts");
        return rc;
    }
    int sum(const uint32_t* nparts, uint32_t nq, uint32_t* out) {
        int rc = launch_sum(nparts, nq);
        return rc ? rc : to_host(out, counts.p, nq);
    }
};

} // namespace

namespace fbh_impl {

// Literal result of a range query (groups + 1 words: the last one holds the
// active bits left-aligned) and its ones count as *nparts per-block partial
// counts in part[].  plan: FBH_PLAN_LITERAL or FBH_PLAN_SCAN.
int index_eval_literal(fbh_index* ix, double lo, int lop, double hi, int rop, int plan,
                       uint32_t* out, uint32_t* part, uint32_t* nparts, hipStream_t s) {
    fbk::QueryArgs q;
    fill_query(ix, q, lo, lop, hi, rop);
    q.out = out;
    q.part = part;
    *nparts = plan == FBH_PLAN_SCAN ? fbk::scan_blocks(ix->groups, ix->anb) : ix->ntiles;
    Locate L;
    locate(ix, lo, lop, hi, rop, L);
    if (plan == FBH_PLAN_SCAN) {
        // part::doScan over every row: integer columns take the bounds
        // converted to the element type (fill_query), floats compare as
        // doubles
        return check(fbk::launch_query_scan(q, dev_type(ix->type), ix->nrows, s), "scan");
    }
    set_sum_ranges(ix, q, L.hit0, L.hit1);
    if (ix->kind == FBH_BIN) {
        if (L.cand0 < L.hit0 && !ix->null_bitmap(L.cand0)) q.cand0 = L.cand0;
        if (L.cand1 > L.hit1 && L.hit1 < ix->nobs && !ix->null_bitmap(L.hit1)) q.cand1 = L.hit1;
    }
    return check(fbk::launch_query_lit(q, dev_type(ix->type), s), "query");
}

} // namespace fbh_impl

using namespace fbh_impl;

namespace {

// Device output of the compressed plan (launch_query_wah): result words and
// chunk index sized for the worst case, tile look-back status, control words.
struct WahBuf {
    DBuf<uint32_t> out, ck, ctl;
    DBuf<uint64_t> status;
    uint32_t tag = 2;          // the tag of the last launch on these buffers
    int alloc(const fbh_index* ix) {
        const uint32_t ns = fbk::wah_status_words(ix->ntiles);
        int rc = out.alloc(ix->groups + 2);
        if (!rc) rc = ck.alloc(ix->groups / fbk::CHUNK_WORDS + 3);
        if (!rc) rc = ctl.alloc(8);
        if (!rc) rc = status.alloc(ns);
        if (!rc) rc = check(fbk::launch_zero_u64(status.p, ns, stream()), "zero");
        if (!rc) rc = check(fbk::launch_zero_u32(ctl.p, 8, stream()), "zero");
        return rc;
    }
    // arguments of the next launch (tags alternate)
    fbk::WahOut args() {
        tag = 3u - tag;
        return fbk::WahOut{out.p, ck.p, status.p, ctl.p, tag};
    }
};

// Compressed plan (sumBins leaving its result compressed, src/index.cpp:
// 6687-6760): the OR set and the edge rows merged tile by tile on the
// device, the canonical WAH of the result written directly; part[] receives
// ntiles per-tile ones.
int launch_compressed(fbh_index* ix, double lo, int lop, double hi, int rop, WahBuf& wb, uint32_t* part,
                      hipStream_t s) {
    fbk::QueryArgs q;
    fill_query(ix, q, lo, lop, hi, rop);
    q.out = nullptr;
    q.part = part;
    Locate L;
    locate(ix, lo, lop, hi, rop, L);
    set_sum_ranges(ix, q, L.hit0, L.hit1);
    if (ix->kind == FBH_BIN) {
        if (L.cand0 < L.hit0 && !ix->null_bitmap(L.cand0)) q.cand0 = L.cand0;
        if (L.cand1 > L.hit1 && L.hit1 < ix->nobs && !ix->null_bitmap(L.hit1)) q.cand1 = L.hit1;
    }
    return check(fbk::launch_query_wah(q, dev_type(ix->type), wb.args(), s), "compressed");
}

// The compressed plan's result as the words of `res` (its previous words are
// released).
int eval_compressed(fbh_index* ix, double lo, int lop, double hi, int rop, fbh_bitvector* res) {
    WahBuf wb;
    DBuf<uint32_t> part;
    int rc = wb.alloc(ix);
    if (!rc) rc = part.alloc(ix->ntiles);
    if (!rc) rc = launch_compressed(ix, lo, lop, hi, rop, wb, part.p, stream());
    if (!rc) rc = check(fbk::launch_wah_edges(wb.out.p, wb.ctl.p, stream()), "edge words");
    uint32_t h[8];
    if (!rc) rc = to_host(h, wb.ctl.p, 8);
    if (rc) return rc;
    dev_free(res->d); dev_free(res->ck);
    res->cap = ix->groups + 2;
    res->d = wb.out.release();
    res->ckcap = ix->groups / fbk::CHUNK_WORDS + 3;
    res->ck = wb.ck.release();
    res->ck_ok = true;
    res->nw = h[1]; res->w0 = h[4]; res->wl = h[5];
    res->nbits = ix->groups * MAXBITS;
    res->nset = 0;
    res->anb = ix->anb;
    res->aval = ix->anb ? (h[2] >> (31 - ix->anb)) : 0;
    return FBH_OK;
}

int plan_heuristic(const fbh_index* ix, double lo, int lop, double hi, int rop);
double index_cost(const fbh_index* ix, double lo, int lop, double hi, int rop);
double scan_cost(const fbh_index* ix);
double two_level_cost(const fbh_index* ix, double lo, int lop, double hi, int rop);
int tl_build(fbh_index* ix, int enc, uint32_t spec);
int tl_read(fbh_index* ix, int enc, const uint8_t* file, size_t n, size_t pos, size_t w);
double tl_query_bytes(const fbh_index* ix, double lo, int lop, double hi, int rop);
int tl_evaluate(fbh_index* ix, double lo, int lop, double hi, int rop, uint64_t* hits, fbh_bitvector** result,
                uint32_t* rids, uint64_t cap);
int tl_batch_counts(fbh_index* ix, uint32_t nq, const double* lo, const int* lop, const double* hi,
                    const int* rop, const int* plans, int plan, uint64_t* hits, int* plans_used, float* gpu_ms);
int tl_batch_prepare(fbh_index* ix, uint32_t nq, const double* lo, const int* lop, const double* hi,
                     const int* rop, const int* plans, int plan, fbh_conj** out);
// One evaluation as a postfix program over terms of the fine and the
// coarse level (operators as the expression kernel takes them).
struct TLProgram {
    int option = FBH_TL_EMPTY;
    uint64_t loads = 0;                      // coarse bitmaps read
    std::vector<fbk::QueryArgs> terms;
    std::vector<int32_t> prog;
};
void two_level_program(const fbh_index* ix, double lo, int lop, double hi, int rop, TLProgram& P);

// Canonical WAH of a literal result (groups words, then the active bits
// left-aligned) as a new bitvector.
int literal_bitvector(uint32_t groups, uint32_t anb, uint32_t* lit, fbh_bitvector** out) {
    fbh_bitvector* res = nullptr;
    int rc = fbh_bv_new(&res);
    if (rc) return rc;
    uint32_t act = 0;
    if (anb && (rc = to_host(&act, lit + groups, 1))) { fbh_bv_free(res); return rc; }
    fbh_bitvector tmp;
    tmp.d = lit; tmp.nw = groups; tmp.nbits = groups * MAXBITS;
    tmp.anb = anb; tmp.aval = anb ? (act >> (31 - anb)) : 0;
    if (groups) {
        if ((rc = to_host(&tmp.w0, lit, 1)) || (rc = to_host(&tmp.wl, lit + groups - 1, 1))) {
            fbh_bv_free(res); return rc;
        }
    }
    // res = empty | tmp  (the canonical kernel runs through the c0/d1 path)
    rc = fbh_bv_set(res, 0, groups * MAXBITS + anb);
    if (!rc) rc = bv_op(res, &tmp, FBH_OR, nullptr);
    tmp.d = nullptr;
    if (tmp.ck) { dev_free(tmp.ck); tmp.ck = nullptr; }
    if (rc) { fbh_bv_free(res); return rc; }
    *out = res;
    return FBH_OK;
}

// The OR of bitmaps [r0, r1) and [r2, r3), complemented when flip, as a new
// bitvector (the literal kernel without an edge check).
int ranges_bitvector(fbh_index* ix, uint32_t r0, uint32_t r1, uint32_t r2, uint32_t r3, uint32_t flip,
                     fbh_bitvector** out) {
    fbk::QueryArgs q;
    fill_query(ix, q, 0.0, FBH_OP_NONE, 0.0, FBH_OP_NONE);
    q.r0 = r0; q.r1 = r1; q.r2 = r2; q.r3 = r3; q.flip = flip;
    DBuf<uint32_t> lit, part;
    int rc;
    if ((rc = lit.alloc(ix->groups + 1)) || (rc = part.alloc(ix->ntiles))) return rc;
    q.out = lit.p; q.part = part.p;
    if ((rc = check(fbk::launch_query_lit(q, dev_type(ix->type), stream()), "estimate"))) return rc;
    return literal_bitvector(ix->groups, ix->anb, lit.p, out);
}

int literal_bitvector(const fbh_index* ix, uint32_t* lit, fbh_bitvector** out) {
    return literal_bitvector(ix->groups, ix->anb, lit, out);
}

// index::sumBins(ib, ie) as a new bitvector.
int sum_bins_bitvector(fbh_index* ix, uint32_t ib, uint32_t ie, fbh_bitvector** out) {
    fbk::QueryArgs q;
    memset(&q, 0, sizeof(q));
    set_sum_ranges(ix, q, ib, ie);
    return ranges_bitvector(ix, q.r0, q.r1, q.r2, q.r3, q.flip, out);
}

// Installs bitmaps given in the serialized index form (src/iapi.h:160-176):
// bitmap b is bms[offsets[b], offsets[b+1]); an empty range is a null
// bitmap, any other is code words ++ [active value when the active bit
// count is not 0] ++ active bit count (bitvector::write(array_t&),
// src/bitvector.cpp:2283).  The words are checked on the host before a
// kernel reads them, so that a damaged index is an error and not an access
// beyond a bitmap: every bitmap has to cover exactly the column's rows.
// The ones of every bitmap come out of the same pass (the rows-per-bitmap
// table of the cost model).
int attach_bitmaps(fbh_index* ix, const int64_t* offsets, uint32_t nobs, const uint32_t* bms,
                   uint64_t nbms) {
    if (offsets[0] < 0) return fail(FBH_ERR_ARG, "negative bitmap offset");
    for (uint32_t b = 0; b < nobs; ++b)
        if (offsets[b + 1] < offsets[b]) return fail(FBH_ERR_ARG, "bitmap offsets are not ascending");
    if ((uint64_t)offsets[nobs] > nbms) return fail(FBH_ERR_ARG, "bitmap offsets beyond the words");
    const uint64_t base = (uint64_t)offsets[0];
    const uint32_t trailer = ix->anb > 0 ? 2u : 1u;
    std::vector<uint64_t> off(nobs + 1);
    std::vector<uint32_t> nw(nobs);
    ix->cumrows.assign(nobs + 1, 0.0);
    for (uint32_t b = 0; b < nobs; ++b) {
        const uint64_t s = (uint64_t)offsets[b], e = (uint64_t)offsets[b + 1];
        off[b] = s - base;
        nw[b] = 0;
        uint64_t ones = 0;
        if (e > s) {
            if (e - s < trailer) return fail(FBH_ERR_ARG, "bitmap " + std::to_string(b) + " is truncated");
            if (bms[e - 1] != ix->anb)
                return fail(FBH_ERR_ARG, "bitmap " + std::to_string(b) + " does not have the column's rows");
            if (ix->anb) {
                const uint32_t av = bms[e - 2];
                if (av >> ix->anb) return fail(FBH_ERR_ARG, "bitmap " + std::to_string(b) + " has a bad active word");
                ones += (uint32_t)__builtin_popcount(av);
            }
            const uint64_t ncode = e - s - trailer;
            if (ncode > 0xFFFFFFFFull) return fail(FBH_ERR_SIZE, "bitmap larger than 2^32 words");
            uint64_t groups = 0;
            for (uint64_t i = s; i < s + ncode; ++i) {
                const uint32_t w = bms[i];
                if (w & 0x80000000u) {
                    const uint32_t n = w & 0x3FFFFFFFu;
                    if (n == 0) return fail(FBH_ERR_ARG, "bitmap " + std::to_string(b) + " has an empty fill");
                    groups += n;
                    if (w & 0x40000000u) ones += (uint64_t)MAXBITS * n;
                } else {
                    groups += 1;
                    ones += (uint32_t)__builtin_popcount(w);
                }
            }
            if (groups != ix->groups)
                return fail(FBH_ERR_ARG, "bitmap " + std::to_string(b) + " does not have the column's rows");
            nw[b] = (uint32_t)ncode;
        }
        ix->cumrows[b + 1] = ix->cumrows[b] + (double)ones;
    }
    off[nobs] = (uint64_t)offsets[nobs] - base;
    const uint64_t total = off[nobs];
    // one word beyond the last bitmap, like the built index
    DBuf<uint32_t> words;
    int rc = words.alloc(total + 1);
    const uint32_t zero = 0;
    if (!rc) rc = to_dev(words.p, bms + base, total);
    if (!rc) rc = to_dev(words.p + total, &zero, 1);
    if (!rc) rc = stream_sync();
    if (rc) return rc;
    ix->nobs = nobs;
    ix->off.swap(off);
    ix->nw.swap(nw);
    ix->nwords = total;
    ix->words = words.release();
    ix->offsets_built = true;      // the key/offset constructors fill offset64
    return FBH_OK;
}

// Keys of an attached index: relic values (sorted), or bin boundaries with
// the per-bin minima and maxima; then the bitmaps and the tables the plans
// read.  The column range of the overlap test comes from the keys, as for
// the built index.
int attach_index(fbh_index* ix, const double* vals, const double* bounds, const double* minv,
                 const double* maxv, uint32_t nobs, const int64_t* offsets, const uint32_t* bms,
                 uint64_t nbms) {
    if (ix->kind == FBH_RELIC) {
        ix->keys.assign(vals, vals + nobs);
        for (uint32_t b = 1; b < nobs; ++b)
            if (!(ix->keys[b - 1] < ix->keys[b])) return fail(FBH_ERR_ARG, "relic values are not ascending");
    } else {
        ix->keys.assign(bounds, bounds + nobs);
        ix->minval.assign(minv, minv + nobs);
        ix->maxval.assign(maxv, maxv + nobs);
        for (uint32_t b = 1; b < nobs; ++b)
            if (!(ix->keys[b - 1] <= ix->keys[b])) return fail(FBH_ERR_ARG, "bin boundaries are not ascending");
    }
    int rc = attach_bitmaps(ix, offsets, nobs, bms, nbms);
    if (rc) return rc;
    if (ix->kind == FBH_RELIC) {
        if (nobs) { ix->cmin = ix->keys.front(); ix->cmax = ix->keys.back(); }
    } else {
        for (uint32_t b = 0; b < nobs; ++b) {
            if (ix->null_bitmap(b)) continue;
            ix->cmin = std::min(ix->cmin, ix->minval[b]);
            ix->cmax = std::max(ix->cmax, ix->maxval[b]);
        }
    }
    return finish_index(ix);
}

// Bytes of an index file, read through a private file mapping.
struct FileMap {
    const uint8_t* p = nullptr;
    size_t n = 0;
    int open(const char* path) {
        const int fd = ::open(path, O_RDONLY);
        if (fd < 0) return fail(FBH_ERR_ARG, std::string("cannot open ") + path);
        struct stat st;
        if (fstat(fd, &st) != 0 || st.st_size < 0) { ::close(fd); return fail(FBH_ERR_ARG, "cannot stat the file"); }
        n = (size_t)st.st_size;
        if (n) {
            void* m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
            if (m == MAP_FAILED) { ::close(fd); n = 0; return fail(FBH_ERR_ARG, "cannot map the file"); }
            p = (const uint8_t*)m;
        }
        ::close(fd);
        return FBH_OK;
    }
    ~FileMap() { if (p) munmap((void*)p, n); }
    template <class T>
    T at(size_t pos) const { T v; memcpy(&v, p + pos, sizeof(T)); return v; }
};

} // namespace

extern "C" {

void fbh_free(void* p) { std::free(p); }

int fbh_index_build_relic(const void* col, int type, uint32_t nrows, fbh_index** out) {
    FBH_API_LOCK();
    if (!out || (nrows && !col)) return fail(FBH_ERR_ARG, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    fbh_index* ix = new fbh_index();
    ix->kind = FBH_RELIC;
    if ((rc = upload_column(ix, col, type, nrows))) { delete ix; return rc; }
    std::vector<double> vals;
    if ((rc = distinct_values(ix, vals))) { delete ix; return rc; }
    // -0.0 and +0.0 are one value of index::mapValues' std::map
    // (src/index.cpp:2552-2611), keyed by the zero that comes first
    for (size_t i = 1; i < vals.size(); ++i) {
        if (vals[i] == 0.0 && vals[i - 1] == 0.0) {
            const float* fc = (const float*)col;
            uint32_t r = 0;
            while (r < nrows && fc[r] != 0.0f) ++r;
            vals[i - 1] = (r < nrows) ? (double)fc[r] : 0.0;
            vals.erase(vals.begin() + i);
            break;
        }
    }
    ix->keys = vals;
    ix->nobs = (uint32_t)vals.size();
    if (!vals.empty()) { ix->cmin = vals.front(); ix->cmax = vals.back(); }
    DBuf<double> dv;
    DBuf<uint32_t> keys;
    if ((rc = dv.alloc(vals.size())) || (rc = keys.alloc(nrows)) || (rc = to_dev(dv.p, vals.data(), vals.size()))) {
        delete ix; return rc;
    }
    rc = check(fbk::launch_keys_relic(ix->col, dev_type(type), nrows, dv.p, ix->nobs, keys.p, stream()),
               "relic keys");
    if (!rc) rc = build_common(ix, keys.p, ix->nobs, false, false, true);
    if (rc) { delete ix; return rc; }
    *out = ix;
    return FBH_OK;
}

int fbh_index_build_relic_card(const uint32_t* ind, uint32_t nrows, uint32_t card, fbh_index** out) {
    FBH_API_LOCK();
    if (!out || (nrows && !ind)) return fail(FBH_ERR_ARG, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    fbh_index* ix = new fbh_index();
    ix->kind = FBH_RELIC;
    if ((rc = upload_column(ix, ind, FBH_COL_UINT32, nrows))) { delete ix; return rc; }
    ix->nobs = card;
    ix->keys.resize(card);
    for (uint32_t i = 0; i < card; ++i) ix->keys[i] = i;
    DBuf<uint32_t> keys;
    if ((rc = keys.alloc(nrows))) { delete ix; return rc; }
    rc = check(fbk::launch_keys_card((const uint32_t*)ix->col, nrows, card, keys.p, stream()), "card keys");
    if (!rc) rc = build_common(ix, keys.p, card, false, false, false);
    if (rc) { delete ix; return rc; }
    ix->offsets_built = true;     // relic(col, card, ind), src/irelic.cpp:186-190
    for (uint32_t i = 0; i < nrows; ++i) {
        ix->cmin = std::min(ix->cmin, (double)ind[i]);
        ix->cmax = std::max(ix->cmax, (double)ind[i]);
    }
    *out = ix;
    return FBH_OK;
}

// ibis::fuzz(col) / ibis::zona(col) (src/ixfuzz.cpp:20-44, src/ixzona.cpp):
// the relic, then coarsen() adds the coarse level.  ncoarse is the index
// spec's "ncoarse=" value (fuzz; 0 or below 5 for the default rule).
static int build_two_level(const void* col, int type, uint32_t nrows, int enc, uint32_t ncoarse, fbh_index** out) {
    fbh_index* ix = nullptr;
    int rc = fbh_index_build_relic(col, type, nrows, &ix);
    if (rc) return rc;
    rc = tl_build(ix, enc, ncoarse);
    if (rc) { delete ix; return rc; }
    *out = ix;
    return FBH_OK;
}

int fbh_index_build_fuzz(const void* col, int type, uint32_t nrows, uint32_t ncoarse, fbh_index** out) {
    FBH_API_LOCK();
    if (!out || (nrows && !col)) return fail(FBH_ERR_ARG, "null argument");
    return build_two_level(col, type, nrows, FBH_ENC_FUZZ, ncoarse, out);
}

int fbh_index_build_zona(const void* col, int type, uint32_t nrows, fbh_index** out) {
    FBH_API_LOCK();
    if (!out || (nrows && !col)) return fail(FBH_ERR_ARG, "null argument");
    return build_two_level(col, type, nrows, FBH_ENC_ZONA, 0, out);
}

int fbh_index_coarse(const fbh_index* ix, int* enc, uint32_t* ncoarse, uint32_t* cbounds, uint32_t cap,
                     int64_t* coffsets, uint32_t ccap, uint32_t* ncbits) {
    FBH_API_LOCK();
    if (!ix) return fail(FBH_ERR_ARG, "null index");
    const uint32_t nc = ix->cbounds.empty() ? 0u : (uint32_t)ix->cbounds.size() - 1;
    const uint32_t nb = ix->coarse ? ix->coarse->nobs : 0u;
    if (enc) *enc = ix->enc;
    if (ncoarse) *ncoarse = nc;
    if (ncbits) *ncbits = nb;
    if (cbounds) for (uint32_t i = 0; i < std::min<uint32_t>(cap, (uint32_t)ix->cbounds.size()); ++i) cbounds[i] = ix->cbounds[i];
    if (coffsets) for (uint32_t i = 0; i < std::min<uint32_t>(ccap, (uint32_t)ix->tcoff.size()); ++i) coffsets[i] = ix->tcoff[i];
    return FBH_OK;
}

int fbh_index_coarse_bitmap(const fbh_index* ix, uint32_t i, fbh_bitvector** out) {
    FBH_API_LOCK();
    if (!ix || !out) return fail(FBH_ERR_ARG, "null argument");
    if (!ix->coarse || i >= ix->coarse->nobs) return fail(FBH_ERR_ARG, "no such coarse bitmap");
    const fbh_index* c = ix->coarse;
    const uint32_t trailer = c->anb ? 2u : 1u;
    if (c->null_bitmap(i)) return fbh_bv_new(out);
    std::vector<uint32_t> w(c->nw[i] + trailer);
    int rc = to_host(w.data(), c->words + c->off[i], w.size());
    return rc ? rc : fbh_bv_from_serial(w.data(), (uint32_t)w.size(), out);
}

int fbh_index_two_level_option(const fbh_index* ix, double lo, int lop, double hi, int rop, int* option,
                               uint64_t* coarse_read) {
    FBH_API_LOCK();
    if (!ix || !option) return fail(FBH_ERR_ARG, "null argument");
    if (!ix->enc) return fail(FBH_ERR_ARG, "not a fuzz or zona index");
    TLProgram P;
    two_level_program(ix, lo, lop, hi, rop, P);
    *option = P.option;
    if (coarse_read) *coarse_read = P.loads;
    return FBH_OK;
}

int fbh_index_build_bin(const void* col, int type, uint32_t nrows, const double* bounds,
                        uint32_t nbounds, fbh_index** out) {
    FBH_API_LOCK();
    if (!out || (nrows && !col) || !bounds || nbounds == 0) return fail(FBH_ERR_ARG, "bad argument");
    int rc = ensure_init();
    if (rc) return rc;
    fbh_index* ix = new fbh_index();
    ix->kind = FBH_BIN;
    if ((rc = upload_column(ix, col, type, nrows))) { delete ix; return rc; }
    ix->keys.assign(bounds, bounds + nbounds);
    ix->nobs = nbounds;
    DBuf<double> db;
    DBuf<uint32_t> keys;
    if ((rc = db.alloc(nbounds)) || (rc = keys.alloc(nrows)) || (rc = to_dev(db.p, bounds, nbounds))) {
        delete ix; return rc;
    }
    rc = check(fbk::launch_keys_bin(ix->col, dev_type(type), nrows, db.p, nbounds, keys.p, stream()),
               "bin keys");
    if (!rc) rc = build_common(ix, keys.p, nbounds, true, true, true);
    if (rc) { delete ix; return rc; }
    for (uint32_t b = 0; b < ix->nobs; ++b) {
        if (ix->null_bitmap(b)) continue;
        ix->cmin = std::min(ix->cmin, ix->minval[b]);
        ix->cmax = std::max(ix->cmax, ix->maxval[b]);
    }
    *out = ix;
    return FBH_OK;
}

int fbh_index_build_bin_nbins(const void* col, int type, uint32_t nrows, uint32_t nbins,
                              fbh_index** out) {
    FBH_API_LOCK();
    if (!out || (nrows && !col) || nbins == 0) return fail(FBH_ERR_ARG, "bad argument");
    if (type == FBH_COL_DOUBLE) return fail(FBH_ERR_ARG, "equal-weight bins over double columns are not supported");
    int rc = ensure_init();
    if (rc) return rc;
    // histogram on the device, boundaries on the host
    fbh_index tmp;
    if ((rc = upload_column(&tmp, col, type, nrows))) return rc;
    std::vector<double> vals;
    std::vector<uint32_t> cnt;
    if ((rc = distinct_values(&tmp, vals, &cnt))) return rc;
    // -0.0 and +0.0 are one histogram entry (std::map<double, uint32_t>)
    for (size_t i = 1; i < vals.size(); ++i) {
        if (vals[i] == vals[i - 1]) {
            cnt[i - 1] += cnt[i];
            vals.erase(vals.begin() + i);
            cnt.erase(cnt.begin() + i);
            --i;
        }
    }
    const bool is_float = type == FBH_COL_FLOAT;
    std::vector<double> bounds;
    // the in-memory column has no computed upper bound (-DBL_MAX)
    fbh_ix::equal_weight_bounds(vals.data(), cnt.data(), (uint32_t)vals.size(), nbins, is_float, -DBL_MAX,
                                bounds);
    if (bounds.empty()) return fail(FBH_ERR_ARG, "empty column");
    return fbh_index_build_bin(col, type, nrows, bounds.data(), (uint32_t)bounds.size(), out);
}

int fbh_index_attach(const void* col, int type, uint32_t nrows, const double* keys, uint64_t nkeys,
                     const int64_t* offsets, uint64_t noffsets, const uint32_t* bms, uint64_t nbms,
                     fbh_index** out) {
    FBH_API_LOCK();
    // the argument checks of fastbit_iapi_attach_full_index
    // (src/iapi.cpp:2037-2044)
    if (!out || (nrows && !col) || !keys || nkeys == 0 || !offsets || noffsets == 0 || !bms || nbms == 0)
        return fail(FBH_ERR_ARG, "bad argument");
    if ((uint64_t)offsets[noffsets - 1] > nbms) return fail(FBH_ERR_ARG, "bitmap offsets beyond the words");
    // column::attachIndex (src/column.cpp:5637-5677): bin when there are
    // two keys per bitmap (minima ++ maxima), relic when there is one
    int kind;
    if (nkeys == 2 * (noffsets - 1)) kind = FBH_BIN;
    else if (nkeys + 1 == noffsets) kind = FBH_RELIC;
    else return fail(FBH_ERR_ARG, "mismatching nkeys and noffsets");
    if (noffsets - 1 > 0xFFFFFFFFull) return fail(FBH_ERR_SIZE, "too many bitmaps");
    const uint32_t nobs = (uint32_t)(noffsets - 1);
    int rc = ensure_init();
    if (rc) return rc;
    fbh_index* ix = new fbh_index();
    ix->kind = kind;
    if ((rc = upload_column(ix, col, type, nrows))) { delete ix; return rc; }
    std::vector<double> bounds;
    if (kind == FBH_BIN) {
        // bin(col, nb, keys, offs, ...) (src/ibin.cpp:358-373): the
        // boundary between two bins is the shortest value between the
        // maximum of one and the minimum of the next; the last is DBL_MAX
        bounds.resize(nobs);
        for (uint32_t j = 0; j + 1 < nobs; ++j) bounds[j] = fbh_ix::compact_value(keys[nobs + j], keys[j + 1]);
        bounds[nobs - 1] = DBL_MAX;
        rc = attach_index(ix, nullptr, bounds.data(), keys, keys + nobs, nobs, offsets, bms, nbms);
    } else {
        rc = attach_index(ix, keys, nullptr, nullptr, nullptr, nobs, offsets, bms, nbms);
    }
    if (rc) { delete ix; return rc; }
    *out = ix;
    return FBH_OK;
}

int fbh_index_read_file(const char* path, const void* col, int type, uint32_t nrows, fbh_index** out) {
    FBH_API_LOCK();
    if (!path || !out || (nrows && !col)) return fail(FBH_ERR_ARG, "null argument");
    FileMap f;
    int rc = f.open(path);
    if (rc) return rc;
    // 8-byte header: "#IBIS", index type, offset width, 0
    // (src/irelic.cpp:704-712, src/ibin.cpp:479-487)
    if (f.n < 16 || memcmp(f.p, "#IBIS", 5) != 0 || (f.p[6] != 4 && f.p[6] != 8) || f.p[7] != 0)
        return fail(FBH_ERR_ARG, std::string(path) + " is not an index file");
    const int itype = f.p[5];
    const size_t w = f.p[6];
    // ibis::index::INDEX_TYPE (src/index.h:91-143)
    constexpr int BINNING = 0, RELIC = 7, FUZZ = 24, ZONA = 25;
    const uint32_t frows = f.at<uint32_t>(8), nobs = f.at<uint32_t>(12);
    size_t obeg, oend, kend;
    std::vector<double> vals, bounds, minv, maxv;
    auto doubles = [&](size_t pos, std::vector<double>& v) {
        v.resize(nobs);
        if (nobs) memcpy(v.data(), f.p + pos, 8ull * nobs);
    };
    if (itype == RELIC || itype == FUZZ || itype == ZONA) {
        // relic::read (src/irelic.cpp:727-749): nrows, nobs, card; the
        // values at byte 24; the nobs+1 bitmap positions after them
        if (f.n < 20) return fail(FBH_ERR_ARG, "truncated index file");
        if (f.at<uint32_t>(16) != nobs) return fail(FBH_ERR_ARG, "relic file with card != nobs");
        obeg = 24 + 8ull * nobs;
        oend = obeg + w * (nobs + 1ull);
        kend = oend;
        if (kend > f.n) return fail(FBH_ERR_ARG, "truncated index file");
        doubles(24, vals);
    } else if (itype == BINNING) {
        // bin::read (src/ibin.cpp:501-540): nrows, nobs, the nobs+1 bitmap
        // positions, then at the next 8-byte boundary bounds, maxval, minval
        obeg = 16;
        oend = obeg + w * (nobs + 1ull);
        const size_t kbeg = 8 * ((7 + oend) / 8);
        kend = kbeg + 24ull * nobs;
        if (kend > f.n) return fail(FBH_ERR_ARG, "truncated index file");
        doubles(kbeg, bounds);
        doubles(kbeg + 8ull * nobs, maxv);
        doubles(kbeg + 16ull * nobs, minv);
    } else {
        return fail(FBH_ERR_ARG, "index type " + std::to_string(itype) +
                                     " is not a relic, fuzz, zona or bin index");
    }
    if (nobs == 0) return fail(FBH_ERR_ARG, "index file without bitmaps");
    if (frows != nrows)
        return fail(FBH_ERR_ARG, "the file indexes " + std::to_string(frows) + " rows, the column has " +
                                     std::to_string(nrows));
    // bitmap positions are byte offsets in the file; the in-memory form
    // counts words from the first bitmap
    std::vector<int64_t> pos(nobs + 1ull);
    for (uint32_t b = 0; b <= nobs; ++b)
        pos[b] = w == 8 ? f.at<int64_t>(obeg + 8ull * b) : (int64_t)f.at<int32_t>(obeg + 4ull * b);
    if (pos[0] < (int64_t)kend || pos[0] % 4 != 0 || (uint64_t)pos[nobs] > f.n)
        return fail(FBH_ERR_ARG, "bitmap positions outside the file");
    std::vector<int64_t> woff(nobs + 1ull);
    for (uint32_t b = 0; b <= nobs; ++b) {
        if (pos[b] < pos[0] || (pos[b] - pos[0]) % 4 != 0) return fail(FBH_ERR_ARG, "bad bitmap position");
        woff[b] = (pos[b] - pos[0]) / 4;
    }
    if ((rc = ensure_init())) return rc;
    fbh_index* ix = new fbh_index();
    ix->kind = itype == BINNING ? FBH_BIN : FBH_RELIC;
    if ((rc = upload_column(ix, col, type, nrows))) { delete ix; return rc; }
    const uint32_t* bms = (const uint32_t*)(f.p + pos[0]);
    rc = attach_index(ix, vals.data(), bounds.data(), minv.data(), maxv.data(), nobs, woff.data(), bms,
                      (uint64_t)woff[nobs]);
    // fuzz / zona: the coarse level follows the last fine bitmap
    if (!rc && (itype == FUZZ || itype == ZONA))
        rc = tl_read(ix, itype == FUZZ ? FBH_ENC_FUZZ : FBH_ENC_ZONA, f.p, f.n, (size_t)pos[nobs], w);
    if (rc) { delete ix; return rc; }
    *out = ix;
    return FBH_OK;
}

int fbh_index_sum_bins(fbh_index* ix, uint32_t ib, uint32_t ie, uint64_t* hits) {
    FBH_API_LOCK();
    if (!ix || !hits) return fail(FBH_ERR_ARG, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    DBuf<uint32_t> lit;
    PartCounts pc;
    if ((rc = lit.alloc(ix->groups + 1)) || (rc = pc.alloc(ix, 1))) return rc;
    fbk::QueryArgs q;
    fill_query(ix, q, 0.0, FBH_OP_NONE, 0.0, FBH_OP_NONE);
    q.out = lit.p;
    q.part = pc.row(0);
    set_sum_ranges(ix, q, ib, ie);
    uint32_t np = ix->ntiles, c = 0;
    rc = check(fbk::launch_query_lit(q, dev_type(ix->type), stream()), "sum bins");
    if (!rc) rc = pc.sum(&np, 1, &c);
    if (rc) return rc;
    *hits = c;
    return FBH_OK;
}

int fbh_index_bounds(const fbh_index* ix, double* bounds, uint32_t cap, uint32_t* n) {
    FBH_API_LOCK();
    if (!ix || !n) return fail(FBH_ERR_ARG, "null argument");
    if (ix->kind != FBH_BIN) return fail(FBH_ERR_ARG, "not a bin index");
    *n = (uint32_t)ix->keys.size();
    if (bounds) memcpy(bounds, ix->keys.data(), sizeof(double) * std::min<size_t>(cap, ix->keys.size()));
    return FBH_OK;
}

void fbh_index_free(fbh_index* ix) {
    FBH_API_LOCK();
    delete ix;
}

int fbh_index_info(const fbh_index* ix, uint32_t* nobs, uint32_t* nrows, uint64_t* nwords) {
    FBH_API_LOCK();
    if (!ix) return fail(FBH_ERR_ARG, "null index");
    if (nobs) *nobs = ix->nobs;
    if (nrows) *nrows = ix->nrows;
    if (nwords) *nwords = ix->nwords;
    return FBH_OK;
}

int fbh_index_deconstruct(const fbh_index* ix, double** keys, uint64_t* nkeys, int64_t** offsets,
                          uint64_t* noffsets, uint32_t** bms, uint64_t* nbms) {
    FBH_API_LOCK();
    if (!ix || !keys || !nkeys || !offsets || !noffsets || !bms || !nbms)
        return fail(FBH_ERR_ARG, "null argument");
    // relic: keys = vals; bin: keys = minval ++ maxval
    // (src/irelic.cpp:644-667, src/ibin.cpp:5836-5860)
    std::vector<double> k;
    if (ix->kind == FBH_RELIC) k = ix->keys;
    else { k = ix->minval; k.insert(k.end(), ix->maxval.begin(), ix->maxval.end()); }
    *nkeys = k.size();
    *keys = (double*)std::malloc(sizeof(double) * (k.size() ? k.size() : 1));
    *noffsets = ix->nobs + 1;
    *offsets = (int64_t*)std::malloc(sizeof(int64_t) * (ix->nobs + 1));
    *nbms = ix->nwords;
    *bms = (uint32_t*)std::malloc(sizeof(uint32_t) * (ix->nwords ? ix->nwords : 1));
    if (!*keys || !*offsets || !*bms) return fail(FBH_ERR_NOMEM, "host allocation failed");
    if (!k.empty()) memcpy(*keys, k.data(), sizeof(double) * k.size());
    for (uint32_t b = 0; b <= ix->nobs; ++b) (*offsets)[b] = (int64_t)ix->off[b];
    return to_host(*bms, ix->words, ix->nwords);
}

int fbh_index_resolve_range(const fbh_index* ix, double lo, int lop, double hi, int rop,
                            uint32_t* cand0, uint32_t* hit0, uint32_t* hit1, uint32_t* cand1) {
    FBH_API_LOCK();
    if (!ix) return fail(FBH_ERR_ARG, "null index");
    Locate L;
    locate_raw(ix, lo, lop, hi, rop, L);
    if (cand0) *cand0 = L.cand0;
    if (hit0) *hit0 = L.hit0;
    if (hit1) *hit1 = L.hit1;
    if (cand1) *cand1 = L.cand1;
    return FBH_OK;
}

int fbh_index_estimate_cost(const fbh_index* ix, double lo, int lop, double hi, int rop,
                            double* icost, double* scost, int* plan) {
    FBH_API_LOCK();
    if (!ix) return fail(FBH_ERR_ARG, "null index");
    const double ic = index_cost(ix, lo, lop, hi, rop);
    const double sc = scan_cost(ix);
    if (icost) *icost = ic;
    if (scost) *scost = sc;
    if (plan) *plan = plan_heuristic(ix, lo, lop, hi, rop);
    return FBH_OK;
}

int fbh_index_set_cost_model(fbh_index* ix, int model) {
    FBH_API_LOCK();
    if (!ix) return fail(FBH_ERR_ARG, "null index");
    if (model != FBH_COST_IAPI && model != FBH_COST_PARTITION) return fail(FBH_ERR_ARG, "unknown cost model");
    ix->cost_model = model;
    return FBH_OK;
}

int fbh_index_evaluate(fbh_index* ix, double lo, int lop, double hi, int rop, int plan,
                       uint64_t* hits, fbh_bitvector** result, int* plan_used) {
    FBH_API_LOCK();
    if (!ix) return fail(FBH_ERR_ARG, "null index");
    if (lop < 0 || lop > 2 || rop < 0 || rop > 2) return fail(FBH_ERR_ARG, "bad operator");
    int rc = ensure_init();
    if (rc) return rc;
    if (plan == FBH_PLAN_AUTO) plan = plan_heuristic(ix, lo, lop, hi, rop);
    if (plan_used) *plan_used = plan;
    if (ix->enc && plan == FBH_PLAN_LITERAL) return tl_evaluate(ix, lo, lop, hi, rop, hits, result, nullptr, 0);
    if (plan == FBH_PLAN_COMPRESSED) {
        fbh_bitvector* res = nullptr;
        if ((rc = fbh_bv_new(&res))) return rc;
        rc = eval_compressed(ix, lo, lop, hi, rop, res);
        uint32_t c = 0;
        if (!rc) rc = fbh_bv_cnt(res, &c);
        if (rc) { fbh_bv_free(res); return rc; }
        if (hits) *hits = c;
        if (result) *result = res; else fbh_bv_free(res);
        return FBH_OK;
    }
    DBuf<uint32_t> lit;
    PartCounts pc;
    if ((rc = lit.alloc(ix->groups + 1)) || (rc = pc.alloc(ix, 1))) return rc;
    uint32_t np = 0;
    rc = index_eval_literal(ix, lo, lop, hi, rop, plan, lit.p, pc.row(0), &np, stream());
    uint32_t c = 0;
    if (!rc) rc = pc.sum(&np, 1, &c);
    if (rc) return rc;
    if (hits) *hits = c;
    if (result) return literal_bitvector(ix, lit.p, result);
    return FBH_OK;
}

int fbh_index_evaluate_rids(fbh_index* ix, double lo, int lop, double hi, int rop, int plan,
                            uint32_t* rids, uint64_t cap, uint64_t* n) {
    FBH_API_LOCK();
    if (!ix || !n) return fail(FBH_ERR_ARG, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    if (plan == FBH_PLAN_AUTO || plan == FBH_PLAN_COMPRESSED) plan = FBH_PLAN_LITERAL;
    if (ix->enc && plan == FBH_PLAN_LITERAL) return tl_evaluate(ix, lo, lop, hi, rop, n, nullptr, rids, cap);
    DBuf<uint32_t> lit;
    PartCounts pc;
    if ((rc = lit.alloc(ix->groups + 1)) || (rc = pc.alloc(ix, 1))) return rc;
    uint32_t np = 0;
    rc = index_eval_literal(ix, lo, lop, hi, rop, plan, lit.p, pc.row(0), &np, stream());
    uint32_t c = 0;
    if (!rc) rc = pc.sum(&np, 1, &c);
    if (rc) return rc;
    *n = c;
    if (!rids || c == 0) return FBH_OK;
    // the active slot is one more literal group (rows groups*31 ..)
    const uint32_t ng = ix->groups + (ix->anb ? 1 : 0);
    DBuf<uint32_t> out, tmp;
    if ((rc = out.alloc(c)) || (rc = tmp.alloc(2ull * ng + fbk::scan_tmp_words(ng) + 4))) return rc;
    rc = check(fbk::launch_rids_lit(lit.p, ng, out.p, c, tmp.p, stream()), "rids");
    if (rc) return rc;
    const uint64_t k = c < cap ? c : cap;
    return to_host(rids, out.p, k);
}

// relic::estimate (src/irelic.h:58-62: evaluate, upper cleared) and
// bin::estimate (src/ibin.cpp:6722-6829).  The cost test reads the bitmap
// sizes as offset64 holds them once sumBins has run: the serialized sizes
// where the index came with offsets, else bitvector::bytes().
int fbh_index_estimate(fbh_index* ix, double lo, int lop, double hi, int rop, fbh_bitvector** lower,
                       fbh_bitvector** upper) {
    FBH_API_LOCK();
    if (!ix || !lower || !upper) return fail(FBH_ERR_ARG, "null argument");
    if (lop < 0 || lop > 2 || rop < 0 || rop > 2) return fail(FBH_ERR_ARG, "bad operator");
    int rc = ensure_init();
    if (rc) return rc;
    fbh_bitvector *lw = nullptr, *up = nullptr;
    auto done = [&](int r) {
        if (r) { fbh_bv_free(lw); fbh_bv_free(up); return r; }
        *lower = lw; *upper = up;
        return FBH_OK;
    };
    if ((rc = fbh_bv_new(&up))) return rc;                   // a cleared upper unless set below
    if (ix->kind == FBH_RELIC) {
        uint64_t h = 0;
        return done(fbh_index_evaluate(ix, lo, lop, hi, rop, FBH_PLAN_LITERAL, &h, &lw, nullptr));
    }
    const uint32_t nobs = ix->nobs;
    if (nobs == 0 || ix->nrows == 0) {
        if ((rc = fbh_bv_new(&lw)) == 0) rc = fbh_bv_set(lw, 0, ix->nrows);
        return done(rc);
    }
    Locate L;
    locate(ix, lo, lop, hi, rop, L);
    const uint32_t c0 = L.cand0, c1 = L.cand1, h0 = L.hit0, h1 = L.hit1;
    double cost = 0.0;
    if (c0 < c1 && c1 < nobs + 1) {
        const std::vector<double>& m = ix->offsets_built ? ix->cumser : ix->cumbytes;
        const int64_t tot = (int64_t)(m[nobs] - m[0]);
        const int64_t mid = (int64_t)(m[c1] - m[c0]);
        cost = (tot >> 1) >= mid ? (double)mid : (double)(tot - mid);
    }
    const double edge = (double)col_esize(ix->type) * (double)ix->nrows / nobs;
    if (h0 > c0 && c1 > h1) cost += edge * 2.0;
    else if (h0 > c0 || c1 > h1) cost += edge;
    if (c0 >= c1) {
        if ((rc = fbh_bv_new(&lw)) == 0) rc = fbh_bv_set(lw, 0, ix->nrows);
        if (!rc) rc = fbh_bv_set(up, 0, ix->nrows);
        return done(rc);
    }
    if (cost > ix->nrows * 0.75) {
        // too costly: no sure hit, every row a candidate
        if ((rc = fbh_bv_new(&lw)) == 0) rc = fbh_bv_set(lw, 0, ix->nrows);
        if (!rc) rc = fbh_bv_set(up, 1, ix->nrows);
        return done(rc);
    }
    if (h0 < h1) {
        rc = sum_bins_bitvector(ix, h0, h1, &lw);
        const bool e0 = c0 < h0 && !ix->null_bitmap(c0);
        const bool e1 = c1 > h1 && h1 < nobs && !ix->null_bitmap(h1);
        if (!rc && (c0 < h0 || (c1 > h1 && h1 < nobs))) {
            // lower | the edge bins: the bins partition the rows, so the
            // straight OR widens to [c0, c1) and the complement's outside
            // set shrinks by the edge bins
            const uint32_t b0 = e0 ? c0 : h0, b1 = e1 ? h1 + 1 : h1;
            const SumPlan sp = sum_plan(ix, h0, h1);
            fbh_bitvector* u = nullptr;
            if (sp.kind == 1) rc = ranges_bitvector(ix, 0, 0, 0, 0, 1, &u);
            else if (sp.kind == 3) rc = ranges_bitvector(ix, 0, b0, b1, nobs, 1, &u);
            else rc = ranges_bitvector(ix, b0, b1, 0, 0, 0, &u);
            if (!rc) { fbh_bv_free(up); up = u; }
        }
        return done(rc);
    }
    if ((rc = fbh_bv_new(&lw)) == 0) rc = fbh_bv_set(lw, 0, ix->nrows);
    fbh_bitvector* u = nullptr;
    if (!rc) rc = sum_bins_bitvector(ix, c0, c1, &u);
    if (!rc) { fbh_bv_free(up); up = u; }
    return done(rc);
}

// fbh_index_evaluate without any device allocation: the literal result goes
// to the caller's device buffer, the per-block partial counts and the count
// to the library's static scratch array (NOALLOC_PARTS + 1 words per GPU).
int fbh_index_evaluate_noalloc(fbh_index* ix, double lo, int lop, double hi, int rop, int plan,
                               uint32_t* dev_out, uint64_t out_words, uint64_t* hits, int* plan_used) {
    FBH_API_LOCK();
    if (!ix || !dev_out || !hits) return fail(FBH_ERR_ARG, "null argument");
    if (lop < 0 || lop > 2 || rop < 0 || rop > 2) return fail(FBH_ERR_ARG, "bad operator");
    if (plan == FBH_PLAN_COMPRESSED)
        return fail(FBH_ERR_ARG, "the compressed plan allocates; use the literal or scan plan");
    if (plan < FBH_PLAN_AUTO || plan > FBH_PLAN_SCAN) return fail(FBH_ERR_ARG, "bad plan");
    int rc = ensure_init();
    if (rc) return rc;
    uint32_t* sc = noalloc_scratch_ptr();
    if (!sc) return fail(FBH_ERR_DEVICE, "no scratch address for this device");
    if (plan == FBH_PLAN_AUTO) {
        plan = plan_heuristic(ix, lo, lop, hi, rop);
        if (plan == FBH_PLAN_COMPRESSED) plan = FBH_PLAN_LITERAL;     // the same bits
    }
    if (out_words < (uint64_t)ix->groups + 1)
        return fail(FBH_ERR_SIZE, "output buffer smaller than the literal result (groups + 1 words)");
    const uint32_t nparts = plan == FBH_PLAN_SCAN ? fbk::scan_blocks(ix->groups, ix->anb) : ix->ntiles;
    if (nparts > NOALLOC_PARTS)
        return fail(FBH_ERR_SIZE, "index beyond the allocation-free limit of 65536 tiles (2^27 groups)");
    uint32_t np = 0;
    rc = index_eval_literal(ix, lo, lop, hi, rop, plan, dev_out, sc, &np, stream());
    if (!rc) rc = check(fbk::launch_sum_parts_n(sc, np, sc + NOALLOC_PARTS, stream()), "sum parts");
    uint32_t c = 0;
    if (!rc) rc = check(hipMemcpyAsync(&c, sc + NOALLOC_PARTS, 4, hipMemcpyDeviceToHost, stream()), "count copy");
    if (!rc) rc = stream_sync();
    if (rc) return rc;
    *hits = c;
    if (plan_used) *plan_used = plan;
    return FBH_OK;
}

int fbh_index_evaluate_batch(fbh_index* ix, uint32_t nq, const double* lo, const int* lop,
                             const double* hi, const int* rop, const int* plans, int plan,
                             uint64_t* hits, int* plans_used, float* gpu_ms) {
    FBH_API_LOCK();
    if (!ix || !lo || !lop || !hi || !rop || !hits) return fail(FBH_ERR_ARG, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    if (ix->enc) {
        if (nq == 0) return FBH_OK;
        return tl_batch_counts(ix, nq, lo, lop, hi, rop, plans, plan, hits, plans_used, gpu_ms);
    }
    // consecutive literal / scan queries rotate over the side streams, so
    // that the last blocks of one query overlap the first of the next (each
    // stream has its own literal result buffer)
    DBuf<uint32_t> lit[NSIDE];
    PartCounts pc;
    for (int k = 0; k < NSIDE; ++k)
        if ((rc = lit[k].alloc(ix->groups + 1))) return rc;
    if ((rc = pc.alloc(ix, nq))) return rc;
    std::vector<uint32_t> np(nq, 0);
    std::vector<int> used(nq);
    WahBuf wbs[NSIDE];
    hipEvent_t e0, e1, ej[NSIDE];
    (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
    for (int k = 0; k < NSIDE; ++k) (void)hipEventCreateWithFlags(&ej[k], hipEventDisableTiming);
    rc = check(hipEventRecord(e0, stream()), "event");
    for (int k = 0; k < NSIDE && !rc; ++k) rc = check(hipStreamWaitEvent(side_stream(k), e0, 0), "stream wait");
    uint32_t nlaunch = 0;
    for (uint32_t i = 0; i < nq && !rc; ++i) {
        int p = plans ? plans[i] : plan;
        if (p == FBH_PLAN_AUTO) p = plan_heuristic(ix, lo[i], lop[i], hi[i], rop[i]);
        used[i] = p;
        const int k = (int)(nlaunch++ % (uint32_t)NSIDE);
        if (p == FBH_PLAN_COMPRESSED) {
            // the compressed result of each side stream in its own buffers
            if (!wbs[k].out.p) rc = wbs[k].alloc(ix);
            np[i] = ix->ntiles;
            if (!rc) rc = launch_compressed(ix, lo[i], lop[i], hi[i], rop[i], wbs[k], pc.row(i), side_stream(k));
        } else {
            rc = index_eval_literal(ix, lo[i], lop[i], hi[i], rop[i], p, lit[k].p, pc.row(i), &np[i],
                                    side_stream(k));
        }
    }
    // join: stream() waits for every side stream whatever happened above
    for (int k = 0; k < NSIDE; ++k) {
        const int jrc = check(hipEventRecord(ej[k], side_stream(k)), "event");
        const int wrc = jrc ? jrc : check(hipStreamWaitEvent(stream(), ej[k], 0), "stream wait");
        if (!rc) rc = wrc;
    }
    std::vector<uint32_t> c(nq);
    if (!rc) rc = pc.launch_sum(np.data(), nq);
    if (!rc) rc = check(hipEventRecord(e1, stream()), "event");
    if (!rc) rc = check(hipEventSynchronize(e1), "event sync");
    if (rc) {
        // a failed join leaves side-stream kernels that may still write lit[]
        // and the partial counts: drain every stream before they are freed
        (void)hipStreamSynchronize(stream());
        for (int k = 0; k < NSIDE; ++k) (void)hipStreamSynchronize(side_stream(k));
    }
    float ms = 0.f;
    if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    for (int k = 0; k < NSIDE; ++k) (void)hipEventDestroy(ej[k]);
    if (rc) return rc;
    if ((rc = to_host(c.data(), pc.counts.p, nq))) return rc;
    for (uint32_t i = 0; i < nq; ++i) hits[i] = c[i];
    if (plans_used) for (uint32_t i = 0; i < nq; ++i) plans_used[i] = used[i];
    if (gpu_ms) *gpu_ms = ms;
    return FBH_OK;
}

// Algorithmic bytes of one range query under a plan: the bitmap words the
// plan reads (OR set, edge bitmaps), the raw values it reads, and the
// literal result it writes.
int fbh_index_query_bytes(const fbh_index* ix, double lo, int lop, double hi, int rop, int plan,
                          double* bytes) {
    FBH_API_LOCK();
    if (!ix || !bytes) return fail(FBH_ERR_ARG, "null argument");
    const double res = 4.0 * (ix->groups + 1);
    const double es = (double)col_esize(ix->type);
    if (plan == FBH_PLAN_SCAN) { *bytes = es * ix->nrows + res; return FBH_OK; }
    if (ix->enc && plan != FBH_PLAN_COMPRESSED) { *bytes = tl_query_bytes(ix, lo, lop, hi, rop); return FBH_OK; }
    Locate L;
    locate(ix, lo, lop, hi, rop, L);
    const SumPlan sp = sum_plan(ix, L.hit0, L.hit1);
    auto words = [&](uint32_t b0, uint32_t b1) {
        return b1 > b0 ? ix->cumser[b1] - ix->cumser[b0] : 0.0;
    };
    double w = 0;
    if (sp.kind == 2) w = words(L.hit0, L.hit1);
    else if (sp.kind == 3) w = words(0, L.hit0) + words(L.hit1, ix->nobs);
    double edge = 0;
    if (ix->kind == FBH_BIN) {
        const double rows_per_bin = ix->nobs ? (double)ix->nrows / ix->nobs : 0.0;
        if (L.cand0 < L.hit0) edge += words(L.cand0, L.cand0 + 1) + es * rows_per_bin;
        if (L.cand1 > L.hit1 && L.hit1 < ix->nobs) edge += words(L.hit1, L.hit1 + 1) + es * rows_per_bin;
    }
    *bytes = w + edge + res;
    return FBH_OK;
}

int fbh_index_column_bytes(const fbh_index* ix, uint64_t* bytes) {
    FBH_API_LOCK();
    if (!ix || !bytes) return fail(FBH_ERR_ARG, "null argument");
    *bytes = (uint64_t)col_esize(ix->type) * ix->nrows;
    return FBH_OK;
}

int fbh_index_plan_features(const fbh_index* ix, uint32_t nq, const double* lo, const int* lop,
                            const double* hi, const int* rop, float* feats) {
    FBH_API_LOCK();
    if (!ix || (nq && (!lo || !lop || !hi || !rop || !feats))) return fail(FBH_ERR_ARG, "null argument");
    const uint32_t nobs = ix->nobs;
    auto words = [&](uint32_t b0, uint32_t b1) {
        b1 = std::min(b1, nobs);
        return b1 > b0 ? (ix->cumser[b1] - ix->cumser[b0]) / 4.0 : 0.0;
    };
    auto rows = [&](uint32_t b0, uint32_t b1) {
        b1 = std::min(b1, nobs);
        return b1 > b0 ? ix->cumrows[b1] - ix->cumrows[b0] : 0.0;
    };
    for (uint32_t i = 0; i < nq; ++i) {
        Locate L;
        locate(ix, lo[i], lop[i], hi[i], rop[i], L);
        const uint32_t h0 = std::min(L.hit0, nobs), h1 = std::min(L.hit1, nobs);
        const SumPlan sp = sum_plan(ix, h0, h1);
        const double win = words(h0, h1), wout = words(0, h0) + words(h1, nobs);
        double wread = sp.kind == 2 ? win : sp.kind == 3 ? wout : 0.0;
        const double na = sp.kind == 2 ? (double)(h1 - h0) : sp.kind == 3 ? (double)(h0 + nobs - h1) : 0.0;
        double erows = 0.0;
        if (ix->kind == FBH_BIN) {
            if (L.cand0 < L.hit0) { erows += rows(L.cand0, L.cand0 + 1); wread += words(L.cand0, L.cand0 + 1); }
            if (L.cand1 > L.hit1 && L.hit1 < nobs) { erows += rows(L.hit1, L.hit1 + 1); wread += words(L.hit1, L.hit1 + 1); }
        }
        float* f = feats + (size_t)i * FBH_NFEATURES;
        f[0] = (float)std::log2(1.0 + win);
        f[1] = (float)std::log2(1.0 + wout);
        f[2] = (float)std::log2(1.0 + wread);
        f[3] = nobs ? (float)((double)(h1 - h0) / nobs) : 0.f;
        f[4] = ix->nrows ? (float)(rows(h0, h1) / ix->nrows) : 0.f;
        f[5] = (float)std::log2(1.0 + erows);
        f[6] = (float)std::log2(1.0 + na);
        f[7] = sp.kind == 3 ? 1.f : 0.f;
        f[8] = (float)std::log2(1.0 + ix->nrows);
        f[9] = (float)std::log2(1.0 + nobs);
        f[10] = is_int_type(ix->type) ? 0.f : 1.f;
        f[11] = (float)col_esize(ix->type);
    }
    return FBH_OK;
}

int fbh_index_time_plans(fbh_index* ix, uint32_t nq, const double* lo, const int* lop, const double* hi,
                         const int* rop, float cap, float* ms, uint64_t* hits) {
    FBH_API_LOCK();
    if (!ix || (nq && (!lo || !lop || !hi || !rop || !ms || !hits))) return fail(FBH_ERR_ARG, "null argument");
    for (uint32_t i = 0; i < nq; ++i)
        if (lop[i] < 0 || lop[i] > 2 || rop[i] < 0 || rop[i] > 2) return fail(FBH_ERR_ARG, "bad operator");
    int rc = ensure_init();
    if (rc || nq == 0) return rc;
    constexpr uint32_t CH = 256;                       // queries per event chunk
    const uint32_t cmax = std::min(nq, CH);
    DBuf<uint32_t> lit;
    PartCounts pc;
    if ((rc = lit.alloc(ix->groups + 1)) || (rc = pc.alloc(ix, cmax))) return rc;
    std::vector<hipEvent_t> ev(cmax + 1);
    for (auto& e : ev)
        if ((rc = check(hipEventCreate(&e), "event"))) return rc;
    WahBuf wb;
    if ((rc = wb.alloc(ix))) return rc;
    std::vector<uint32_t> np(cmax), cnt(cmax);
    // one launch sequence of a plan for query q (part row i)
    auto launch = [&](int plan, uint32_t q, uint32_t i) {
        if (plan == FBH_PLAN_COMPRESSED) {
            np[i] = ix->ntiles;
            return launch_compressed(ix, lo[q], lop[q], hi[q], rop[q], wb, pc.row(i), stream());
        }
        return index_eval_literal(ix, lo[q], lop[q], hi[q], rop[q], plan, lit.p, pc.row(i), &np[i], stream());
    };
    for (int plan = FBH_PLAN_LITERAL; plan <= FBH_PLAN_SCAN && !rc; ++plan) {
        // untimed warm-up of the plan's kernels
        for (uint32_t i = 0; i < std::min(nq, 4u) && !rc; ++i) rc = launch(plan, i, 0);
        if (!rc) rc = stream_sync();
        for (uint32_t c0 = 0; c0 < nq && !rc; c0 += CH) {
            const uint32_t k = std::min(CH, nq - c0);
            rc = check(hipEventRecord(ev[0], stream()), "event");
            for (uint32_t i = 0; i < k && !rc; ++i) {
                rc = launch(plan, c0 + i, i);
                if (!rc) rc = check(hipEventRecord(ev[i + 1], stream()), "event");
            }
            if (!rc) rc = pc.launch_sum(np.data(), k);
            if (!rc) rc = stream_sync();
            if (!rc) rc = to_host(cnt.data(), pc.counts.p, k);
            for (uint32_t i = 0; i < k && !rc; ++i) {
                float t = 0.f;
                rc = check(hipEventElapsedTime(&t, ev[i], ev[i + 1]), "elapsed");
                ms[3 * (size_t)(c0 + i) + (plan - 1)] = t;
                hits[3 * (size_t)(c0 + i) + (plan - 1)] = cnt[i];
            }
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    (void)cap;
    return rc;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Conjunctions.
struct fbh_conj {
    uint32_t nq = 0, nterms = 0, ntiles = 0, groups = 0, anb = 0;
    fbk::QueryArgs* dterms = nullptr;    // device term descriptors
    uint32_t* dqoff = nullptr;
    uint32_t* dnp = nullptr;             // ntiles per query (for the sums)
    uint32_t* parts = nullptr;           // nq * ntiles
    uint32_t* counts = nullptr;          // nq
    int32_t* dprog = nullptr;            // postfix programs (expression batches)
    int32_t* dskip = nullptr;            // short-circuit targets of the programs
    uint32_t* dpoff = nullptr;
    uint32_t depth = 0;                  // deepest stack of the programs
    std::vector<int> plans;              // plan of every term (caller's order)
    std::vector<uint32_t> order;         // evaluation order: order[k] = caller's index
    int nt = 256;                        // threads per block
    int single_type = -1;                // >= 0: single-term queries on one column type
    int has_scan = 0;                    // 0 no term runs the scan plan, 1 some, 2 all
    bool lit_only = false;               // literal terms only, no edge bitmaps
    uint32_t blocks = 0;                 // persistent grid
    uint32_t tpi = 1;                    // tiles per work item
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~fbh_conj() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        dev_free(dterms); dev_free(dqoff); dev_free(dnp); dev_free(parts); dev_free(counts);
        dev_free(dprog); dev_free(dskip); dev_free(dpoff);
    }
};

namespace {

int cu_count() {
    static int n = 0;
    if (!n) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    }
    return n;
}

// Device descriptor of one range term under the given plan.
void term_args(const fbh_term& t, int p, fbk::QueryArgs& q) {
    fill_query(t.ix, q, t.lo, t.lop, t.hi, t.rop);
    q.type = dev_type(t.ix->type);
    q.plan = p == FBH_PLAN_SCAN ? fbk::PLAN_SCAN_D : fbk::PLAN_LIT_D;
    Locate L;
    locate(t.ix, t.lo, t.lop, t.hi, t.rop, L);
    set_sum_ranges(t.ix, q, L.hit0, L.hit1);
    if (t.ix->kind == FBH_BIN) {
        if (L.cand0 < L.hit0 && !t.ix->null_bitmap(L.cand0)) q.cand0 = L.cand0;
        if (L.cand1 > L.hit1 && L.hit1 < t.ix->nobs && !t.ix->null_bitmap(L.hit1)) q.cand1 = L.hit1;
    }
}

int conj_launch(fbh_conj* c, hipStream_t s);
int conj_capture(fbh_conj* c);

// Term descriptors, query offsets and programs to the device, then the
// launch sequence captured.
int conj_upload(fbh_conj* c, const std::vector<fbk::QueryArgs>& h, const uint32_t* qoff, const int32_t* prog,
                const uint32_t* poff, uint32_t nprog) {
    const uint32_t nq = c->nq;
    std::vector<uint32_t> np(nq, c->ntiles);
    c->dterms = (fbk::QueryArgs*)dev_alloc(sizeof(fbk::QueryArgs) * c->nterms);
    c->dqoff = (uint32_t*)dev_alloc(4ull * (nq + 1));
    c->dnp = (uint32_t*)dev_alloc(4ull * nq);
    c->parts = (uint32_t*)dev_alloc(4ull * nq * c->ntiles);
    c->counts = (uint32_t*)dev_alloc(4ull * nq);
    if (!c->dterms || !c->dqoff || !c->dnp || !c->parts || !c->counts)
        return fail(FBH_ERR_NOMEM, "device allocation failed");
    if (prog) {
        c->dprog = (int32_t*)dev_alloc(4ull * nprog);
        c->dpoff = (uint32_t*)dev_alloc(4ull * (nq + 1));
        if (!c->dprog || !c->dpoff) return fail(FBH_ERR_NOMEM, "device allocation failed");
    }
    int rc = to_dev(c->dterms, h.data(), c->nterms);
    if (!rc) rc = to_dev(c->dqoff, qoff, nq + 1);
    if (!rc) rc = to_dev(c->dnp, np.data(), nq);
    if (!rc && prog) rc = to_dev(c->dprog, prog, nprog);
    if (!rc && prog) rc = to_dev(c->dpoff, poff, nq + 1);
    if (!rc) rc = stream_sync();
    c->lit_only = true;
    for (uint32_t k = 0; k < c->nterms; ++k)
        c->lit_only = c->lit_only && h[k].plan == fbk::PLAN_LIT_D && h[k].cand0 == fbk::NOKEY &&
                      h[k].cand1 == fbk::NOKEY;
    // single-term queries on one column type run the persistent range
    // kernel: 256 threads, 2 tiles per work item, 9 blocks per CU (the
    // tuner's choice on 2^27 rows with 10,000 bins; all resident by LDS)
    bool single = !prog;
    for (uint32_t q = 0; single && q < nq; ++q) single = qoff[q + 1] == qoff[q] + 1;
    for (uint32_t k = 0; single && k < c->nterms; ++k) single = h[k].type == h[0].type;
    if (single) {
        c->single_type = h[0].type;
        uint32_t ns = 0;
        for (uint32_t k = 0; k < c->nterms; ++k) ns += h[k].plan == fbk::PLAN_SCAN_D ? 1u : 0u;
        c->has_scan = ns == 0 ? 0 : ns == c->nterms ? 2 : 1;
        c->nt = 256;
        c->tpi = 2;
        c->blocks = 9u * (uint32_t)cu_count();
    } else {
        c->blocks = 4u * (uint32_t)cu_count();
    }
    if (!rc) rc = conj_capture(c);
    return rc;
}

// Kernel arguments of queries [q0, q0 + nq) of the batch (part / roff rows
// start at query q0).
fbk::ConjArgs conj_args(const fbh_conj* c, uint32_t q0, uint32_t nq) {
    fbk::ConjArgs a;
    a.terms = c->dterms; a.qoff = c->dqoff + q0; a.nq = nq; a.ntiles = c->ntiles;
    a.groups = c->groups; a.anb = c->anb; a.tpi = c->tpi; a.part = c->parts + (size_t)q0 * c->ntiles;
    a.out = nullptr;
    a.prog = c->dprog; a.poff = c->dpoff ? c->dpoff + q0 : nullptr; a.depth = c->depth; a.skip = c->dskip;
    a.rids = nullptr; a.rbase = nullptr; a.roff = nullptr;
    a.lit_only = c->lit_only ? 1u : 0u;
    return a;
}

// The batch kernel over the queries of `a`.  The scan-only form of the range
// kernel counts in registers; passes that need each tile's result in LDS
// (row ids) take the mixed form.
int conj_kernel(const fbh_conj* c, const fbk::ConjArgs& a, hipStream_t s) {
    if (c->single_type >= 0) {
        const int scan = (a.rids && c->has_scan == 2) ? 1 : c->has_scan;
        return check(fbk::launch_range(a, c->single_type, scan, c->nt, c->blocks, s), "range batch");
    }
    return check(fbk::launch_conj(a, c->nt, c->blocks, s), "conjunctions");
}

int conj_launch(fbh_conj* c, hipStream_t s) {
    int rc = conj_kernel(c, conj_args(c, 0, c->nq), s);
    if (!rc) rc = check(fbk::launch_sum_parts(c->parts, c->ntiles, c->dnp, c->nq, c->counts, s), "sum parts");
    return rc;
}

// (Re)captures the launch sequence of the batch in a graph on the library
// stream; a replay is then one graph launch.
int conj_capture(fbh_conj* c) {
    if (c->exec) { (void)hipGraphExecDestroy(c->exec); c->exec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    int rc = check(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) return rc;
    const int lrc = conj_launch(c, stream());
    rc = check(hipStreamEndCapture(stream(), &c->graph), "end capture");
    if (lrc) return lrc;
    if (rc) return rc;
    return check(hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0), "graph instantiate");
}

int conj_run(fbh_conj* c, float* ms) {
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
    int rc = check(hipEventRecord(e0, stream()), "event");
    if (!rc) rc = check(hipGraphLaunch(c->exec, stream()), "graph launch");
    if (!rc) rc = check(hipEventRecord(e1, stream()), "event");
    if (!rc) rc = check(hipEventSynchronize(e1), "event sync");
    float t = 0.f;
    if (!rc) (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    return rc;
}

} // namespace

extern "C" {

int fbh_conj_prepare(const fbh_term* terms, const uint32_t* qoff, uint32_t nq, int plan, fbh_conj** out) {
    FBH_API_LOCK();
    if (!terms || !qoff || !out || nq == 0) return fail(FBH_ERR_ARG, "bad argument");
    if (plan < FBH_PLAN_AUTO || plan > FBH_PLAN_SCAN) return fail(FBH_ERR_ARG, "bad plan");
    int rc = ensure_init();
    if (rc) return rc;
    for (uint32_t q = 0; q < nq; ++q)
        if (qoff[q + 1] <= qoff[q]) return fail(FBH_ERR_ARG, "a query without terms");
    const uint32_t nterms = qoff[nq];
    const fbh_index* ix0 = terms[0].ix;
    if (!ix0) return fail(FBH_ERR_ARG, "null index");
    for (uint32_t k = 0; k < nterms; ++k) {
        const fbh_term& t = terms[k];
        if (!t.ix || t.ix->nrows != ix0->nrows) return fail(FBH_ERR_ARG, "terms over different row counts");
        if (t.lop < 0 || t.lop > 2 || t.rop < 0 || t.rop > 2) return fail(FBH_ERR_ARG, "bad operator");
    }
    fbh_conj* c = new fbh_conj();
    c->nq = nq; c->nterms = nterms; c->ntiles = ix0->ntiles; c->groups = ix0->groups; c->anb = ix0->anb;
    // per query: the terms cheapest first (qExpr::reorder by query::weight,
    // src/qExpr.cpp:1022; here by the bytes each term's plan reads)
    c->plans.resize(nterms);
    c->order.resize(nterms);
    for (uint32_t q = 0; q < nq; ++q) {
        std::vector<std::pair<double, uint32_t>> w;
        for (uint32_t k = qoff[q]; k < qoff[q + 1]; ++k) {
            const fbh_term& t = terms[k];
            int p = plan == FBH_PLAN_AUTO ? plan_heuristic(t.ix, t.lo, t.lop, t.hi, t.rop) : plan;
            if (p == FBH_PLAN_COMPRESSED) p = FBH_PLAN_LITERAL;
            double bytes = 0;
            fbh_index_query_bytes(t.ix, t.lo, t.lop, t.hi, t.rop, p, &bytes);
            w.push_back({bytes, k});
            c->plans[k] = p;
        }
        std::stable_sort(w.begin(), w.end());
        for (uint32_t i = 0; i < (uint32_t)w.size(); ++i) c->order[qoff[q] + i] = w[i].second;
    }
    std::vector<fbk::QueryArgs> h(nterms);
    for (uint32_t k = 0; k < nterms; ++k) term_args(terms[c->order[k]], c->plans[c->order[k]], h[k]);
    rc = conj_upload(c, h, qoff, nullptr, nullptr, 0);
    if (rc) { delete c; return rc; }
    *out = c;
    return FBH_OK;
}

int fbh_expr_prepare(const fbh_term* terms, const uint32_t* toff, const int32_t* prog, const uint32_t* poff,
                     uint32_t nq, int plan, fbh_conj** out) {
    FBH_API_LOCK();
    if (!terms || !toff || !prog || !poff || !out || nq == 0) return fail(FBH_ERR_ARG, "bad argument");
    if (plan < FBH_PLAN_AUTO || plan > FBH_PLAN_SCAN) return fail(FBH_ERR_ARG, "bad plan");
    int rc = ensure_init();
    if (rc) return rc;
    const uint32_t nterms = toff[nq];
    if (nterms == 0) return fail(FBH_ERR_ARG, "no terms");
    const fbh_index* ix0 = terms[0].ix;
    if (!ix0) return fail(FBH_ERR_ARG, "null index");
    for (uint32_t k = 0; k < nterms; ++k) {
        const fbh_term& t = terms[k];
        if (!t.ix || t.ix->nrows != ix0->nrows) return fail(FBH_ERR_ARG, "terms over different row counts");
        if (t.lop < 0 || t.lop > 2 || t.rop < 0 || t.rop > 2) return fail(FBH_ERR_ARG, "bad operator");
    }
    // every program leaves exactly one result and stays within the stack;
    // its term items refer to the query's own terms
    std::vector<int32_t> aprog(poff[nq]), skip(poff[nq], -1);
    int maxsp = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        if (toff[q + 1] < toff[q] || poff[q + 1] <= poff[q]) return fail(FBH_ERR_ARG, "bad offsets");
        int sp = 0;
        uint32_t start[fbk::EXPR_DEPTH];      // first item of each stacked result's subtree
        for (uint32_t p = poff[q]; p < poff[q + 1]; ++p) {
            const int32_t op = prog[p];
            if (op >= 0) {
                if ((uint32_t)op >= toff[q + 1] - toff[q]) return fail(FBH_ERR_ARG, "term index out of range");
                if (sp >= fbk::EXPR_DEPTH) return fail(FBH_ERR_ARG, "expression deeper than the stack");
                start[sp++] = p;
                maxsp = std::max(maxsp, sp);
                aprog[p] = (int32_t)(toff[q] + (uint32_t)op);
            } else if (op == FBH_EXPR_NOT) {
                if (sp < 1) return fail(FBH_ERR_ARG, "NOT of an empty stack");
                aprog[p] = op;
            } else if (op <= FBH_EXPR_AND && op >= FBH_EXPR_MINUS) {
                if (sp < 2) return fail(FBH_ERR_ARG, "operator without two operands");
                // AND / MINUS with an empty left operand give the empty
                // left operand: its right operand need not run
                if (op == FBH_EXPR_AND || op == FBH_EXPR_MINUS) skip[start[sp - 1]] = (int32_t)(p + 1);
                --sp;
                aprog[p] = op;
            } else {
                return fail(FBH_ERR_ARG, "unknown operator");
            }
        }
        if (sp != 1) return fail(FBH_ERR_ARG, "expression does not leave one result");
    }
    fbh_conj* c = new fbh_conj();
    c->nq = nq; c->nterms = nterms; c->ntiles = ix0->ntiles; c->groups = ix0->groups; c->anb = ix0->anb;
    c->depth = (uint32_t)maxsp;
    c->plans.resize(nterms);
    c->order.resize(nterms);
    std::vector<fbk::QueryArgs> h(nterms);
    for (uint32_t k = 0; k < nterms; ++k) {
        const fbh_term& t = terms[k];
        int p = plan == FBH_PLAN_AUTO ? plan_heuristic(t.ix, t.lo, t.lop, t.hi, t.rop) : plan;
        if (p == FBH_PLAN_COMPRESSED) p = FBH_PLAN_LITERAL;
        c->plans[k] = p;
        c->order[k] = k;
        term_args(t, p, h[k]);
    }
    // a program that is one chain of ANDs runs its terms cheapest first, as
    // a conjunction does (qExpr::reorder by weight, src/qExpr.cpp:1022; AND
    // commutes, so the rows do not change and the chain keeps its shape)
    for (uint32_t q = 0; q < nq; ++q) {
        const uint32_t p0 = poff[q], p1 = poff[q + 1];
        bool chain = (p1 - p0) >= 3 && (p1 - p0) % 2 == 1 && prog[p0] >= 0;
        for (uint32_t p = p0 + 1; chain && p < p1; p += 2)
            chain = prog[p] >= 0 && prog[p + 1] == FBH_EXPR_AND;
        if (!chain) continue;
        std::vector<std::pair<double, int32_t>> w;
        for (uint32_t p = p0; p < p1; p += (p == p0 ? 1u : 2u)) {
            const fbh_term& t = terms[aprog[p]];
            double bytes = 0;
            fbh_index_query_bytes(t.ix, t.lo, t.lop, t.hi, t.rop, c->plans[aprog[p]], &bytes);
            w.push_back({bytes, aprog[p]});
        }
        std::stable_sort(w.begin(), w.end(),
                         [](const std::pair<double, int32_t>& x, const std::pair<double, int32_t>& y) {
                             return x.first < y.first;
                         });
        uint32_t i = 0;
        for (uint32_t p = p0; p < p1; p += (p == p0 ? 1u : 2u)) aprog[p] = w[i++].second;
    }
    // the skip table first: conj_upload captures the launch with its address
    c->dskip = (int32_t*)dev_alloc(4ull * poff[nq]);
    rc = c->dskip ? to_dev(c->dskip, skip.data(), poff[nq]) : fail(FBH_ERR_NOMEM, "device allocation failed");
    if (!rc) rc = conj_upload(c, h, toff, aprog.data(), poff, poff[nq]);
    if (rc) { delete c; return rc; }
    *out = c;
    return FBH_OK;
}

int fbh_conj_run(fbh_conj* c, uint64_t* hits, float* gpu_ms) {
    FBH_API_LOCK();
    if (!c) return fail(FBH_ERR_ARG, "null batch");
    int rc = conj_run(c, gpu_ms);
    if (rc || !hits) return rc;
    std::vector<uint32_t> h(c->nq);
    if ((rc = to_host(h.data(), c->counts, c->nq))) return rc;
    for (uint32_t i = 0; i < c->nq; ++i) hits[i] = h[i];
    return FBH_OK;
}

int fbh_conj_tune(fbh_conj* c, int* nt, uint32_t* blocks, uint32_t* tpi, float* best_ms) {
    FBH_API_LOCK();
    if (!c) return fail(FBH_ERR_ARG, "null batch");
    static const int NTS[] = {128, 256, 512};
    static const int NTS1[] = {192, 256, 256};        // single-term batches
    static const uint32_t TPIS[] = {1, 2, 4};
    static const uint32_t BPC[] = {2, 4, 8};
    static const uint32_t BPC1[] = {4, 8, 9};
    int bnt = c->nt;
    uint32_t bblocks = c->blocks, btpi = c->tpi;
    float best = 0.f;
    int rc = conj_run(c, &best);        // the current configuration, warmed
    if (!rc) rc = conj_run(c, &best);
    for (int a = 0; a < 3 && !rc; ++a)
        for (int b = 0; b < 3 && !rc; ++b)
            for (int d = 0; d < 3 && !rc; ++d) {
                if (c->single_type >= 0 && a == 2) continue;
                c->nt = c->single_type >= 0 ? NTS1[a] : NTS[a];
                c->tpi = TPIS[b];
                c->blocks = (c->single_type >= 0 ? BPC1[d] : BPC[d]) * (uint32_t)cu_count();
                rc = conj_capture(c);
                float ms = 0.f;
                if (!rc) rc = conj_run(c, &ms);
                if (!rc && ms < best) { best = ms; bnt = c->nt; bblocks = c->blocks; btpi = c->tpi; }
            }
    c->nt = bnt; c->blocks = bblocks; c->tpi = btpi;
    if (!rc) rc = conj_capture(c);
    if (nt) *nt = bnt;
    if (blocks) *blocks = bblocks;
    if (tpi) *tpi = btpi;
    if (best_ms) *best_ms = best;
    return rc;
}

int fbh_conj_rids(fbh_conj* c, uint64_t* offsets, uint32_t* rids, uint64_t cap) {
    FBH_API_LOCK();
    if (!c || !offsets) return fail(FBH_ERR_ARG, "null argument");
    // the counts, and each tile's ones in parts
    int rc = conj_run(c, nullptr);
    if (rc) return rc;
    const uint32_t nq = c->nq;
    std::vector<uint32_t> cnt(nq);
    if ((rc = to_host(cnt.data(), c->counts, nq))) return rc;
    offsets[0] = 0;
    for (uint32_t q = 0; q < nq; ++q) offsets[q + 1] = offsets[q] + cnt[q];
    if (!rids || cap == 0 || offsets[nq] == 0) return FBH_OK;
    // each tile's offset within its query
    DBuf<uint32_t> roff;
    DBuf<uint64_t> rbase;
    if ((rc = roff.alloc((size_t)nq * c->ntiles)) || (rc = rbase.alloc(nq))) return rc;
    rc = check(fbk::launch_scan_parts(c->parts, c->ntiles, c->ntiles, nq, roff.p, stream()), "scan parts");
    if (rc) return rc;
    // queries in chunks of at most 2^28 ids on the device (1 GiB), at
    // least one query each; ids past `cap` are not written
    constexpr uint64_t CHUNK_IDS = 1ull << 28;
    std::vector<uint64_t> base(nq);
    uint32_t q0 = 0;
    while (q0 < nq && offsets[q0] < cap) {
        uint32_t q1 = q0 + 1;
        while (q1 < nq && offsets[q1 + 1] - offsets[q0] <= CHUNK_IDS && offsets[q1] < cap) ++q1;
        const uint64_t n = offsets[q1] - offsets[q0];
        if (n) {
            for (uint32_t q = q0; q < q1; ++q) base[q] = offsets[q] - offsets[q0];
            DBuf<uint32_t> out;
            if ((rc = out.alloc(n))) return rc;
            if ((rc = to_dev(rbase.p + q0, base.data() + q0, q1 - q0))) return rc;
            fbk::ConjArgs a = conj_args(c, q0, q1 - q0);
            a.rids = out.p; a.rbase = rbase.p + q0; a.roff = roff.p + (size_t)q0 * c->ntiles;
            if ((rc = conj_kernel(c, a, stream()))) return rc;
            const uint64_t k = std::min(n, cap - offsets[q0]);
            if ((rc = to_host(rids + offsets[q0], out.p, k))) return rc;
        }
        q0 = q1;
    }
    return FBH_OK;
}

int fbh_conj_result(fbh_conj* c, uint32_t q, fbh_bitvector** out) {
    FBH_API_LOCK();
    if (!c || !out) return fail(FBH_ERR_ARG, "null argument");
    if (q >= c->nq) return fail(FBH_ERR_ARG, "query index out of range");
    DBuf<uint32_t> lit, part;
    int rc;
    if ((rc = lit.alloc(c->groups + 1)) || (rc = part.alloc(c->ntiles))) return rc;
    fbk::ConjArgs a = conj_args(c, q, 1);
    a.part = part.p; a.out = lit.p;
    if ((rc = conj_kernel(c, a, stream()))) return rc;
    return literal_bitvector(c->groups, c->anb, lit.p, out);
}

int fbh_conj_plans(const fbh_conj* c, int* plans) {
    FBH_API_LOCK();
    if (!c || !plans) return fail(FBH_ERR_ARG, "null argument");
    for (size_t k = 0; k < c->plans.size(); ++k) plans[k] = c->plans[k];
    return FBH_OK;
}

void fbh_conj_free(fbh_conj* c) {
    FBH_API_LOCK();
    delete c;
}

int fbh_index_batch_prepare(fbh_index* ix, uint32_t nq, const double* lo, const int* lop, const double* hi,
                            const int* rop, const int* plans, int plan, fbh_conj** out) {
    FBH_API_LOCK();
    if (!ix || !lo || !lop || !hi || !rop || !out || nq == 0) return fail(FBH_ERR_ARG, "bad argument");
    if (plan < FBH_PLAN_AUTO || plan > FBH_PLAN_SCAN) return fail(FBH_ERR_ARG, "bad plan");
    for (uint32_t i = 0; i < nq; ++i) {
        if (plans && (plans[i] < FBH_PLAN_AUTO || plans[i] > FBH_PLAN_SCAN)) return fail(FBH_ERR_ARG, "bad plan");
        if (lop[i] < 0 || lop[i] > 2 || rop[i] < 0 || rop[i] > 2) return fail(FBH_ERR_ARG, "bad operator");
    }
    if (ix->enc) {
        int rc = ensure_init();
        return rc ? rc : tl_batch_prepare(ix, nq, lo, lop, hi, rop, plans, plan, out);
    }
    std::vector<fbh_term> t(nq);
    std::vector<uint32_t> qoff(nq + 1);
    for (uint32_t i = 0; i < nq; ++i) {
        if (plans && (plans[i] < FBH_PLAN_AUTO || plans[i] > FBH_PLAN_SCAN)) return fail(FBH_ERR_ARG, "bad plan");
        t[i] = fbh_term{ix, lo[i], lop[i], hi[i], rop[i]};
        qoff[i] = i;
    }
    qoff[nq] = nq;
    if (!plans) return fbh_conj_prepare(t.data(), qoff.data(), nq, plan, out);
    // per-query plans: prepare under the literal plan, then switch the scan
    // terms (the device descriptors carry the plan)
    int rc = fbh_conj_prepare(t.data(), qoff.data(), nq, FBH_PLAN_LITERAL, out);
    if (rc) return rc;
    fbh_conj* c = *out;
    std::vector<fbk::QueryArgs> h(nq);
    uint32_t ns = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        int p = plans[i] == FBH_PLAN_AUTO ? plan_heuristic(ix, lo[i], lop[i], hi[i], rop[i]) : plans[i];
        if (p != FBH_PLAN_SCAN) p = FBH_PLAN_LITERAL;
        c->plans[i] = p;
        ns += p == FBH_PLAN_SCAN ? 1u : 0u;
        term_args(t[i], p, h[i]);
    }
    c->has_scan = ns == 0 ? 0 : ns == nq ? 2 : 1;
    rc = to_dev(c->dterms, h.data(), nq);
    if (!rc) rc = stream_sync();
    if (!rc) rc = conj_capture(c);
    if (rc) { delete c; *out = nullptr; }
    return rc;
}

int fbh_conj_evaluate(const fbh_term* terms, uint32_t nterms, int plan, uint64_t* hits) {
    FBH_API_LOCK();
    if (!terms || nterms == 0 || !hits) return fail(FBH_ERR_ARG, "bad argument");
    const uint32_t qoff[2] = {0, nterms};
    fbh_conj* c = nullptr;
    int rc = fbh_conj_prepare(terms, qoff, 1, plan, &c);
    if (rc) return rc;
    rc = fbh_conj_run(c, hits, nullptr);
    fbh_conj_free(c);
    return rc;
}

} // extern "C"

namespace {

// relic::estimateCost / bin::estimateCost (src/irelic.cpp:1857-1902,
// src/ibin.cpp:6531-6569).  offset64 (serialized bytes per bitmap) is
// there when the index was read from a file or built by relic(col, card,
// ind); the in-memory constructors leave it empty, and then relic sums
// bitvector::bytes() over the hit bitmaps and bin keeps only the edge term.
double index_cost(const fbh_index* ix, double lo, int lop, double hi, int rop) {
    if (ix->enc) return two_level_cost(ix, lo, lop, hi, rop);
    const bool offsets = ix->cost_model == FBH_COST_PARTITION || ix->offsets_built;
    const uint32_t nobs = ix->nobs;
    auto ser = [&](uint32_t b0, uint32_t b1) {
        b1 = std::min(b1, nobs);
        return b1 > b0 ? ix->cumser[b1] - ix->cumser[b0] : 0.0;
    };
    Locate L;
    locate(ix, lo, lop, hi, rop, L);
    if (ix->kind == FBH_RELIC) {
        const uint32_t h0 = L.hit0, h1 = L.hit1;
        if (h0 >= h1) return 0.0;
        if (offsets) {
            if (h1 > h0 + 1) {
                const double tot = ser(0, nobs), mid = ser(h0, h1);
                return std::floor(tot / 2) >= mid ? mid : tot - mid;
            }
            return 0.5 * ser(h0, h1);
        }
        if (h1 > h0 + 1) {
            const uint32_t e = std::min(h1, nobs);
            return e > h0 ? ix->cumbytes[e] - ix->cumbytes[h0] : 0.0;
        }
        return h0 < nobs ? 0.5 * ix->bytes(h0) : 0.0;
    }
    double r = 0;
    uint32_t cand0 = 0, cand1 = nobs, hit0 = nobs, hit1 = 0;
    if (offsets) {
        cand0 = L.cand0; cand1 = L.cand1; hit0 = L.hit0; hit1 = L.hit1;
        if (cand0 < cand1 && cand1 <= nobs) {
            const double tot = ser(0, nobs), mid = ser(cand0, cand1);
            r = std::floor(tot / 2) >= mid ? mid : tot - mid;
        }
    }
    if (hit0 > cand0 || hit1 < cand1) {
        // column::elementSize() * nrows / nobs in unsigned 32-bit arithmetic
        const uint32_t esn = (uint32_t)col_esize(ix->type) * ix->nrows;
        if (nobs > 0) {
            const double ccheck = (double)(esn / nobs);
            r += (hit0 > cand0 && hit1 < cand1 && hit0 <= hit1) ? 2.0 * ccheck : ccheck;
        } else {
            r += (double)esn;
        }
    }
    return r;
}

// part::countPages over a mask of n ones (src/part.cpp:2616-2678): the
// indexSet walk sees [0, 31g) as one range, then the active word's rows.
uint32_t full_mask_pages(uint32_t n, uint32_t es) {
    if (n == 0) return 0;
    const uint32_t wpp = 4096u / es;
    const uint32_t g = n / MAXBITS, anb = n % MAXBITS;
    uint32_t res = 0, last = 0;
    if (g > 0) { res += (MAXBITS * g) / wpp; last = MAXBITS * g; }
    if (anb > 0) {
        const uint32_t p0 = (MAXBITS * g) / wpp;
        res += (last < p0 * wpp) ? 1u : 0u;
        last = n - 1;
        res += (last / wpp > p0) ? 1u : 0u;
    }
    return res == 0 ? 1u : res;
}

// The cost the index cost is compared with: bord::column::evaluateRange
// (arrays in memory, src/bord.cpp:5903-5912) uses the index below
// nrows/2 + 999 bytes; column::evaluateRange (partitions on disk,
// src/column.cpp:6102-6111) below the scan's page bytes at 4096-byte pages.
double scan_cost(const fbh_index* ix) {
    if (ix->cost_model == FBH_COST_PARTITION)
        return 4096.0 * full_mask_pages(ix->nrows, (uint32_t)col_esize(ix->type)) +
               8.0 * ix->nrows / 4096.0;
    return ix->nrows * 0.5 + 999.0;
}

// qContinuousRange::overlap (src/qExpr.cpp:1673-1803) of lo OP x OP hi
// with the column's [min, max]
bool overlap(const fbh_index* ix, double lo, int lop, double hi, int rop) {
    if (!(ix->cmin <= ix->cmax)) return true;
    if (lop == FBH_OP_LT && !(lo < ix->cmax)) return false;
    if (lop == FBH_OP_LE && !(lo <= ix->cmax)) return false;
    if (rop == FBH_OP_LT) return ix->cmin < hi;
    if (rop == FBH_OP_LE) return ix->cmin <= hi;
    return lop != FBH_OP_NONE;
}

// The ibis heuristic mapped onto the three plans: the index when its cost
// is below the scan cost, else the raw scan (src/bord.cpp:5903-5912 or
// src/column.cpp:6102-6111 by the cost model); a range that misses the
// column's [min, max] under the partition model is answered without
// either (src/column.cpp:6063-6066) -- the literal plan returns its empty
// answer.  With the index, sumBins' priority-queue branch
// (src/index.cpp:6673) is a compressed merge, its decompressed-result
// branches are the literal plan.
int plan_heuristic(const fbh_index* ix, double lo, int lop, double hi, int rop) {
    if (ix->cost_model == FBH_COST_PARTITION && !overlap(ix, lo, lop, hi, rop)) return FBH_PLAN_LITERAL;
    if (!(index_cost(ix, lo, lop, hi, rop) < scan_cost(ix))) return FBH_PLAN_SCAN;
    if (ix->enc) return FBH_PLAN_LITERAL;      // fuzz / zona: the option program
    Locate L;
    locate(ix, lo, lop, hi, rop, L);
    const SumPlan sp = sum_plan(ix, L.hit0, L.hit1);
    if (sp.kind < 2) return FBH_PLAN_LITERAL;
    uint32_t ib = sp.kind == 2 ? L.hit0 : 0, ie = sp.kind == 2 ? L.hit1 : L.hit0;
    double bytes;
    uint32_t na;
    if (sp.kind == 2) {
        bytes = ix->cumbytes[ie] - ix->cumbytes[ib];
        na = ie - ib;
    } else {
        bytes = ix->cumbytes[L.hit0] + ix->cumbytes[ix->nobs] - ix->cumbytes[L.hit1];
        na = L.hit0 + (ix->nobs - L.hit1);
    }
    const double uncomp = (double)(ix->nrows * 4u / 31u);
    const double sum2 = (na >= 2) ? ix->bytes(ib) + ix->bytes(ib + 1 < ix->nobs ? ib + 1 : ib) : bytes;
    if (na > 2 && sum2 < uncomp && bytes * (double)na * na <= std::log(2.0) * uncomp)
        return FBH_PLAN_COMPRESSED;
    return FBH_PLAN_LITERAL;
}

} // namespace

// ---------------------------------------------------------------------------
// Two-level encodings over a relic: ibis::fuzz (the coarse level interval
// encoded: bitmap i holds coarse bins [i, i + ceil(nc/2))) and ibis::zona
// (equality encoded: bitmap i holds coarse bin i), src/irelic.h:540-712.
// The fine level is the relic itself; the coarse bitmaps form an index of
// their own (words, tiled copy), and an evaluation is a short postfix
// program over terms of both levels run by the expression kernel.
namespace {

// offset64 of the fine level in bytes: the file's positions where the index
// came with them, else bitvector::bytes() of every bitmap, as coarsen fills
// it (src/ixfuzz.cpp:166-180, src/ixzona.cpp:130-142)
std::vector<int64_t> tl_fine_offsets(const fbh_index* ix) {
    std::vector<int64_t> off(ix->nobs + 1, 0);
    for (uint32_t b = 0; b < ix->nobs; ++b)
        off[b + 1] = off[b] + (int64_t)(ix->offsets_built ? 4.0 * (double)(ix->off[b + 1] - ix->off[b]) : ix->bytes(b));
    return off;
}

// array_t::find (src/array_t.cpp:507-528): the first position whose value
// is not less than v
template <class T>
uint32_t tl_find(const std::vector<T>& a, T v) {
    return (uint32_t)(std::lower_bound(a.begin(), a.end(), v) - a.begin());
}

// The coarse bin boundaries: the fine bitmaps in ncoarse groups of nearly
// equal bytes, each boundary the fine position nearest its byte target
// (src/ixfuzz.cpp:215-231, src/ixzona.cpp:146-161; fuzz holds the target in
// 32 bits)
std::vector<uint32_t> tl_bounds(const std::vector<int64_t>& off, uint32_t ncoarse, bool target32) {
    const uint32_t nbits = (uint32_t)off.size() - 1;
    std::vector<uint32_t> cb(ncoarse + 1, 0);
    for (uint32_t i = 1; i < ncoarse; ++i) {
        const int64_t base = off[cb[i - 1]];
        int64_t t = base + (off.back() - base) / (int64_t)(ncoarse - i + 1);
        if (target32) t = (int32_t)t;
        uint32_t k = tl_find(off, t);
        if (k > cb[i - 1] + 1 && off[k] - t > t - off[k - 1]) --k;
        else if (k <= cb[i - 1]) k = cb[i - 1] + 1;
        cb[i] = k;
    }
    cb[ncoarse] = nbits;
    for (uint32_t i = ncoarse - 1; i > 0 && cb[i + 1] < cb[i]; --i) cb[i] = cb[i + 1] - 1;
    return cb;
}

// Number of coarse bins fuzz::coarsen picks (src/ixfuzz.cpp:182-210): the
// "ncoarse=" value when above 4, else the count that balances the coarse
// level's size against the bytes a query reads; 0 for none
uint32_t tl_fuzz_ncoarse(const fbh_index* ix, const std::vector<int64_t>& off, uint32_t spec) {
    uint32_t nc = spec > 4 ? spec : 0;
    const uint32_t nrows = ix->nrows;
    if (nc < 5u && off.back() > off[0] + (int64_t)(nrows / 31u)) {
        const int wm1 = 31;                                       // bits per literal word
        const long sf = (long)((off.back() - off[0]) / 4);        // fine-level words
        nc = (unsigned)(wm1 * sf / (std::sqrt(2.0) * nrows));
        const unsigned ncmax = (unsigned)std::sqrt(2.0 * ix->nobs);
        if (nc < ncmax) {
            const double obj1 = (sf + (nc + 1 - std::ceil(0.5 * nc)) * nrows / wm1) * (sf * 0.5 / nc + 2.0 * nrows / wm1);
            const double obj2 = (sf + (nc + 2 - std::ceil(0.5 * nc + 0.5)) * nrows / wm1) *
                                (sf * 0.5 / (nc + 1.0) + 2.0 * nrows / wm1);
            nc += (obj2 < obj1) ? 1u : 0u;
        } else {
            nc = ncmax;
        }
    }
    if (nc < 5 || nc >= ix->nobs) return 0;
    return nc;
}

// The coarse bitmaps of an index from their serialized words on the host
// (file layout): an index of their own over the same rows.
int tl_attach(fbh_index* ix, int enc, std::vector<uint32_t> cbounds, const std::vector<int64_t>& coff_bytes,
              const std::vector<int64_t>& woff, const uint32_t* words) {
    fbh_index* c = new fbh_index();
    c->kind = FBH_RELIC;
    c->type = ix->type;
    c->nrows = ix->nrows; c->groups = ix->groups; c->anb = ix->anb;
    const uint32_t ncb = (uint32_t)woff.size() - 1;
    int rc = attach_bitmaps(c, woff.data(), ncb, words, (uint64_t)woff.back());
    if (!rc) rc = finish_index(c);
    if (rc) { delete c; return rc; }
    delete ix->coarse;
    ix->coarse = c;
    ix->enc = enc;
    ix->cbounds = std::move(cbounds);
    ix->tcoff = coff_bytes;
    return FBH_OK;
}

// fuzz::coarsen / zona::coarsen (src/ixfuzz.cpp:161-263, src/ixzona.cpp:
// 124-182): the coarse bins, then each coarse bitmap as the canonical WAH
// of an OR of consecutive fine bitmaps (the literal kernel, then the
// canonical compression), its bytes() into coffset64.
int tl_build(fbh_index* ix, int enc, uint32_t spec) {
    ix->enc = enc;
    if (ix->nobs < 32) return FBH_OK;                     // no coarse level (nor offsets) below 32 values
    ix->toff = tl_fine_offsets(ix);
    const uint32_t nc = enc == FBH_ENC_FUZZ ? tl_fuzz_ncoarse(ix, ix->toff, spec) : 11u;
    if (nc == 0) return FBH_OK;
    std::vector<uint32_t> cb = tl_bounds(ix->toff, nc, enc == FBH_ENC_FUZZ);
    const uint32_t nc2 = (nc + 1) / 2;
    const uint32_t ncb = enc == FBH_ENC_FUZZ ? nc - nc2 + 1 : nc;
    const uint32_t trailer = ix->anb ? 2u : 1u;
    std::vector<fbh_bitvector*> bvs(ncb, nullptr);
    std::vector<int64_t> woff(ncb + 1, 0), coff(ncb + 1, 0);
    int rc = FBH_OK;
    for (uint32_t i = 0; i < ncb && !rc; ++i) {
        const uint32_t a = cb[i], b = enc == FBH_ENC_FUZZ ? cb[i + nc2] : cb[i + 1];
        rc = ranges_bitvector(ix, a, b, 0, 0, 0, &bvs[i]);
        if (!rc) {
            woff[i + 1] = woff[i] + bvs[i]->nw + trailer;
            coff[i + 1] = coff[i] + 4 * (int64_t)bvs[i]->nw + 40;      // bitvector::bytes()
        }
    }
    // the serialized coarse bitmaps, assembled on the host
    std::vector<uint32_t> words(rc ? 0 : (size_t)woff[ncb]);
    for (uint32_t i = 0; i < ncb && !rc; ++i) {
        fbh_bitvector* v = bvs[i];
        uint32_t* w = words.data() + woff[i];
        if (v->nw) rc = to_host(w, v->d, v->nw);
        if (ix->anb) w[v->nw] = v->aval;
        w[v->nw + trailer - 1] = ix->anb;
    }
    for (fbh_bitvector* v : bvs) if (v) fbh_bv_free(v);
    if (rc) return rc;
    return tl_attach(ix, enc, std::move(cb), coff, woff, words.data());
}


struct TLBuilder {
    const fbh_index* ix;
    TLProgram& P;
    const std::vector<int64_t>& off;
    const std::vector<int64_t>& coff;
    const std::vector<uint32_t>& cb;
    TLBuilder(const fbh_index* x, TLProgram& p)
        : ix(x), P(p), off(x->toff), coff(x->tcoff), cb(x->cbounds) {}
    uint32_t nobs() const { return ix->nobs; }
    uint32_t ncb() const { return ix->coarse ? ix->coarse->nobs : 0u; }

    void push(const fbk::QueryArgs& q) {
        P.prog.push_back((int32_t)P.terms.size());
        P.terms.push_back(q);
    }
    void op(int32_t o) { P.prog.push_back(o); }
    fbk::QueryArgs base(const fbh_index* lvl) const {
        fbk::QueryArgs q;
        fill_query(ix, q, 0.0, FBH_OP_NONE, 0.0, FBH_OP_NONE);
        q.tw = lvl->tw; q.TW = lvl->TW; q.nobs = lvl->nobs;
        q.type = dev_type(ix->type);
        q.plan = fbk::PLAN_LIT_D;
        return q;
    }
    void none() { push(base(ix)); }
    void all() { fbk::QueryArgs q = base(ix); q.flip = 1; push(q); }
    // fine bitmaps [a, b) straight
    void fine(uint32_t a, uint32_t b) {
        fbk::QueryArgs q = base(ix);
        q.r0 = a; q.r1 = std::min(b, nobs());
        push(q);
    }
    // index::sumBins(a, b): the range or the complement of the rest
    void sum_bins(uint32_t a, uint32_t b) {
        fbk::QueryArgs q = base(ix);
        set_sum_ranges(ix, q, a, b);
        push(q);
    }
    // index::addBins(a, b, res) (src/index.cpp:5537-5580): OR of [a, b),
    // every row when that is the whole fine level
    void add_bins(uint32_t a, uint32_t b) {
        b = std::min(b, nobs());
        if (a >= b) return;
        if (a == 0 && b == nobs()) all(); else fine(a, b);
        op(fbk::EXPR_OR);
    }
    void minus_sum(uint32_t a, uint32_t b) { sum_bins(a, b); op(fbk::EXPR_MINUS); }
    // coarse bitmap i; false when it is null
    bool have(uint32_t i) const { return i < ncb() && !ix->coarse->null_bitmap(i); }
    void coarse(uint32_t i) {
        P.loads |= (i < 64) ? (1ull << i) : 0ull;
        fbk::QueryArgs q = base(ix->coarse);
        q.r0 = i; q.r1 = i + 1;
        push(q);
    }
    // OR of coarse bitmaps [a, b), or with flip the rows outside [a0, a1) and
    // [b0, b1)
    void coarse_ranges(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1, bool flip) {
        fbk::QueryArgs q = base(ix->coarse);
        q.r0 = a0; q.r1 = std::max(a0, a1); q.r2 = b0; q.r3 = std::max(b0, b1); q.flip = flip ? 1u : 0u;
        for (uint32_t i = a0; i < a1 && i < 64; ++i) if (have(i)) P.loads |= 1ull << i;
        for (uint32_t i = b0; i < b1 && i < 64; ++i) if (have(i)) P.loads |= 1ull << i;
        push(q);
    }

    long fine_cost(uint32_t h0, uint32_t h1) const {
        const long in = (long)(off[h1] - off[h0]);
        const long out = (long)((off.back() - off[h1]) + (off[h0] - off[0]));
        return in <= out ? in : out;
    }
    long csize(uint32_t i) const { return (long)(coff[i + 1] - coff[i]); }
    // fuzz::coarseEstimate (src/ixfuzz.cpp:626-674)
    long fuzz_ccost(uint32_t lo, uint32_t hi) const {
        const uint32_t mid = (uint32_t)cb.size() / 2;
        if (lo >= cb.size() || lo >= hi) return 0;
        if (hi > mid) {
            long c = csize(hi - mid);
            if (lo > hi - mid) c += lo >= mid ? csize(lo - mid) : csize(lo);
            else if (lo < hi - mid) c += csize(lo);
            return c;
        }
        if (hi < mid) return csize(lo) + csize(hi);
        return csize(0) + (lo > 0 ? csize(lo) : 0);
    }
    // fuzz::coarseEvaluate (src/ixfuzz.cpp:676-765): coarse bins [lo, hi)
    // from one or two interval bitmaps
    void fuzz_coarse(uint32_t lo, uint32_t hi) {
        const uint32_t mid = (uint32_t)cb.size() / 2;
        if (lo >= cb.size() || lo >= hi) { none(); return; }
        if (lo + 1 == hi) {
            const uint32_t a = hi < ncb() ? lo : hi - mid, b = hi < ncb() ? hi : lo - mid;
            if (!have(a)) { none(); return; }
            coarse(a);
            if (have(b)) { coarse(b); op(fbk::EXPR_MINUS); }
            return;
        }
        if (hi > mid) {
            if (have(hi - mid)) coarse(hi - mid); else none();
            if (lo > hi - mid) {
                if (lo >= mid) {
                    if (have(lo - mid)) { coarse(lo - mid); op(fbk::EXPR_MINUS); }
                } else if (have(lo)) {
                    coarse(lo); op(fbk::EXPR_AND);
                } else {
                    none(); op(fbk::EXPR_AND);
                }
            } else if (lo < hi - mid) {
                if (have(lo)) { coarse(lo); op(fbk::EXPR_OR); }
            }
            return;
        }
        if (hi < mid) {
            if (!have(lo)) { none(); return; }
            coarse(lo);
            if (have(hi)) { coarse(hi); op(fbk::EXPR_MINUS); }
            return;
        }
        if (have(0)) coarse(0); else none();
        if (lo > 0 && have(lo)) { coarse(lo); op(fbk::EXPR_AND); }
    }
    // zona: the coarse part of a pair (src/ixzona.cpp:886-1010): bitmaps
    // [a, b) when they hold at most half the coarse bytes, else the rows
    // outside [lo_end, hi_begin) complemented
    long zona_part(uint32_t a, uint32_t b) const {
        const long all = (long)(coff.back() - coff[0]);
        const long t = (long)(coff[b] - coff[a]);
        return t + t <= all ? t : all - t;
    }
    void zona_coarse(uint32_t a, uint32_t b, bool lo_cond, uint32_t lo_end, bool hi_cond, uint32_t hi_begin,
                     uint32_t ncoarse) {
        const long all = (long)(coff.back() - coff[0]);
        const long t = (long)(coff[b] - coff[a]);
        if (t + t <= all) coarse_ranges(a, b, 0, 0, false);
        else coarse_ranges(0, lo_cond ? lo_end : 0, hi_cond ? hi_begin : 0, hi_cond ? ncoarse : 0, true);
    }
};

// fuzz::evaluate / zona::evaluate (src/ixfuzz.cpp:872-1095, src/ixzona.cpp:
// 713-1017) as a program: the option, its coarse part and its fine edges.
void two_level_program(const fbh_index* ix, double lo, int lop, double hi, int rop, TLProgram& P) {
    TLBuilder B(ix, P);
    P = TLProgram();
    const uint32_t nobs = ix->nobs;
    Locate L;
    locate(ix, lo, lop, hi, rop, L);
    const uint32_t h0 = L.hit0, h1 = L.hit1;
    const bool fuzz = ix->enc == FBH_ENC_FUZZ;
    if (nobs == 0 || h1 <= h0 || (fuzz && h0 >= nobs)) { B.none(); return; }
    if (h0 == 0 && h1 >= nobs) { P.option = FBH_TL_ALL; B.all(); return; }
    if (h0 + 1 == h1) { P.option = FBH_TL_EQUAL; B.fine(h0, h1); return; }
    const auto& cb = ix->cbounds;
    const auto& off = ix->toff;
    const auto& coff = ix->tcoff;
    const uint32_t ncoarse = fuzz ? (uint32_t)cb.size() - 1 : std::min<uint32_t>(B.ncb(), (uint32_t)cb.size() - 1);
    if (h0 + 3 >= h1 || ncoarse == 0 || coff.size() != B.ncb() + 1) {
        P.option = FBH_TL_FINE; B.sum_bins(h0, h1); return;
    }
    const uint32_t c0 = tl_find(cb, h0), c1 = tl_find(cb, h1);
    auto O = [&](uint32_t i) { return (long)off[i]; };
    int opt;
    if (fuzz) {
        const long finec = B.fine_cost(h0, h1);
        if (c0 >= c1) {
            const long t = B.fuzz_ccost(c1 - 1, c1) + (O(h0) - O(cb[c1 - 1])) + (O(cb[c1]) - O(h1));
            if (finec <= (long)(0.99 * t)) { P.option = FBH_TL_FINE; B.sum_bins(h0, h1); return; }
            P.option = FBH_TL_SAME_BIN;
            B.fuzz_coarse(c1 - 1, c1);
            if (h0 > cb[c1 - 1]) B.minus_sum(cb[c1 - 1], h0);
            if (cb[c1] > h1) B.minus_sum(h1, cb[c1]);
            return;
        }
        // options 2-5 (src/ixfuzz.cpp:943-1011); option 4's coarse cost
        // starts at c0-1 here, where estimateCost starts it at c0
        opt = 2;
        long cost = B.fuzz_ccost(c0, c1 - 1) + (O(cb[c0]) - O(h0)) + (O(h1) - O(cb[c1 - 1]));
        long t;
        if (c0 > 0) {
            t = B.fuzz_ccost(c0 - 1, c1 - 1) + (O(h0) - O(cb[c0 - 1])) + (O(h1) - O(cb[c1 - 1]));
            if (t < cost) { cost = t; opt = 3; }
        }
        t = B.fuzz_ccost(c0 > 0 ? c0 - 1 : 0, c1) + (O(cb[c0]) - O(h0)) + (O(cb[c1]) - O(h1));
        if (t < cost) { cost = t; opt = 4; }
        if (c0 > 0) {
            t = B.fuzz_ccost(c0 - 1, c1) + (O(h0) - O(cb[c0 - 1])) + (O(cb[c1]) - O(h1));
            if (t < cost) { cost = t; opt = 5; }
        }
        if (cost > (long)(0.99 * finec)) opt = 1;
        P.option = opt;
        switch (opt) {
        case 2:
            B.fuzz_coarse(c0, c1 - 1);
            if (h0 < cb[c0]) B.add_bins(h0, cb[c0]);
            if (cb[c1 - 1] < h1) B.add_bins(cb[c1 - 1], h1);
            break;
        case 3:
            B.fuzz_coarse(c0 - 1, c1 - 1);
            if (cb[c0 - 1] < h0) B.minus_sum(cb[c0 - 1], h0);
            if (cb[c1 - 1] < h1) B.add_bins(cb[c1 - 1], h1);
            break;
        case 4:
            B.fuzz_coarse(c0, c1);
            if (h0 < cb[c0]) B.add_bins(h0, cb[c0]);
            if (cb[c1] > h1) B.minus_sum(h1, cb[c1]);
            break;
        case 5:
            B.fuzz_coarse(c0 - 1, c1);
            if (h0 > cb[c0 - 1]) B.minus_sum(cb[c0 - 1], h0);
            if (cb[c1] > h1) B.minus_sum(h1, cb[c1]);
            break;
        default:
            B.sum_bins(h0, h1);
        }
        return;
    }
    if (c0 >= c1) {
        const long fine = O(h1) - O(h0);
        const long t = (long)(coff[c0] - coff[c0 - 1]) + (O(cb[c0]) - O(cb[c0 - 1])) - fine;
        if ((long)(0.99 * fine) <= t) { P.option = FBH_TL_FINE; B.sum_bins(h0, h1); return; }
        P.option = FBH_TL_SAME_BIN;
        if (B.have(c0 - 1)) B.coarse(c0 - 1); else B.all();
        if (h0 > cb[c0 - 1]) B.minus_sum(cb[c0 - 1], h0);
        if (h1 < cb[c0]) B.minus_sum(h1, cb[c0]);
        return;
    }
    // five pairs (src/ixzona.cpp:803-882)
    opt = 2;
    long cost = B.zona_part(c0, c1 - 1) + (O(cb[c0]) - O(h0)) + (O(h1) - O(cb[c1 - 1]));
    long t;
    if (c0 > 0) {
        t = B.zona_part(c0 - 1, c1 - 1) + (O(h0) - O(cb[c0 - 1])) + (O(h1) - O(cb[c1 - 1]));
        if (t < cost) { cost = t; opt = 3; }
    }
    t = B.zona_part(c0, c1) + (O(cb[c0]) - O(h0)) + (O(cb[c1]) - O(h1));
    if (t < cost) { cost = t; opt = 4; }
    if (c0 > 0) {
        t = B.zona_part(c0 - 1, c1) + (O(h0) - O(cb[c0 - 1])) + (O(cb[c1]) - O(h1));
        if (t < cost) { cost = t; opt = 5; }
    }
    if (cost > (long)(0.99 * B.fine_cost(h0, h1))) opt = 1;
    P.option = opt;
    switch (opt) {
    case 2:
        if (c0 < c1 - 1) B.zona_coarse(c0, c1 - 1, c0 > 0, c0, c1 <= ncoarse, c1 - 1, ncoarse);
        else B.none();
        if (h0 < cb[c0]) B.add_bins(h0, cb[c0]);
        if (cb[c1 - 1] < h1) B.add_bins(cb[c1 - 1], h1);
        break;
    case 3:
        B.zona_coarse(c0 - 1, c1 - 1, c0 > 1, c0 - 1, c1 <= ncoarse, c1 - 1, ncoarse);
        if (cb[c0 - 1] < h0) B.minus_sum(cb[c0 - 1], h0);
        if (cb[c1 - 1] < h1) B.add_bins(cb[c1 - 1], h1);
        break;
    case 4:
        B.zona_coarse(c0, c1, c0 > 0, c0, c1 < ncoarse, c1, ncoarse);
        if (h0 < cb[c0]) B.add_bins(h0, cb[c0]);
        if (cb[c1] > h1) B.minus_sum(h1, cb[c1]);
        break;
    case 5:
        B.zona_coarse(c0 - 1, c1, c0 > 1, c0 - 1, c1 < ncoarse, c1, ncoarse);
        if (h0 > cb[c0 - 1]) B.minus_sum(cb[c0 - 1], h0);
        if (cb[c1] > h1) B.minus_sum(h1, cb[c1]);
        break;
    default:
        B.sum_bins(h0, h1);
    }
}

// fuzz::estimateCost / zona::estimateCost (src/ixfuzz.cpp:767-870,
// src/ixzona.cpp:580-711)
double two_level_cost(const fbh_index* ix, double lo, int lop, double hi, int rop) {
    TLProgram P;
    TLBuilder B(ix, P);
    const uint32_t nobs = ix->nobs;
    if (nobs == 0 || ix->toff.size() != nobs + 1) return (double)col_esize(ix->type) * ix->nrows;
    Locate L;
    locate(ix, lo, lop, hi, rop, L);
    const uint32_t h0 = L.hit0, h1 = L.hit1;
    const auto& cb = ix->cbounds;
    const auto& off = ix->toff;
    const auto& coff = ix->tcoff;
    auto O = [&](uint32_t i) { return (long)off[i]; };
    const bool have_coff = coff.size() == B.ncb() + 1 && B.ncb() > 0;
    if (ix->enc == FBH_ENC_FUZZ) {
        if (h1 <= h0 || h0 >= nobs) return 0.0;
        if (h0 == 0 && h1 >= nobs) return 0.0;
        const uint32_t ncoarse = cb.empty() ? 0u : (uint32_t)cb.size() - 1;
        const long fine = B.fine_cost(h0, h1);
        if (h0 + 3 >= h1 || ncoarse == 0 || !have_coff) return (double)fine;
        const uint32_t c0 = tl_find(cb, h0), c1 = tl_find(cb, h1);
        if (c0 >= c1) {
            const long t = B.fuzz_ccost(c1 - 1, c1) + (O(h0) - O(cb[c1 - 1])) + (O(cb[c1]) - O(h1));
            return (double)(t / 100 >= fine / 99 ? fine : t);
        }
        long cost = B.fuzz_ccost(c0, c1 - 1) + (O(cb[c0]) - O(h0)) + (O(h1) - O(cb[c1 - 1]));
        if (c0 > 0)
            cost = std::min(cost, B.fuzz_ccost(c0 - 1, c1 - 1) + (O(h0) - O(cb[c0 - 1])) + (O(h1) - O(cb[c1 - 1])));
        cost = std::min(cost, B.fuzz_ccost(c0, c1) + (O(cb[c0]) - O(h0)) + (O(cb[c1]) - O(h1)));
        if (c0 > 0)
            cost = std::min(cost, B.fuzz_ccost(c0 - 1, c1) + (O(h0) - O(cb[c0 - 1])) + (O(cb[c1]) - O(h1)));
        return (double)(cost / 100 >= fine / 99 ? fine : cost);
    }
    if (h1 <= h0) return 0.0;
    if (h0 == 0 && h1 >= nobs) return 0.0;
    const long fine = B.fine_cost(h0, h1);
    const uint32_t ncoarse = cb.empty() ? 0u : std::min<uint32_t>(B.ncb(), (uint32_t)cb.size() - 1);
    if (h0 + 1 == h1) return (double)fine;
    if (h0 + 3 >= h1 || !have_coff) return (double)fine;
    const uint32_t c0 = tl_find(cb, h0), c1 = tl_find(cb, h1);
    if (c0 >= c1) {
        const long t = (long)(coff[c0] - coff[c0 - 1]) + (O(cb[c0]) - O(cb[c0 - 1])) - fine;
        return (double)((long)(0.99 * fine) >= t ? t : fine);
    }
    if (c1 < ncoarse && cb[c1] == h1 && cb[c0] == h0) return (double)B.zona_part(c0, c1);
    long cost = B.zona_part(c0, c1 - 1) + (O(cb[c0]) - O(h0)) + (O(h1) - O(cb[c1 - 1]));
    if (c0 > 0)
        cost = std::min(cost, B.zona_part(c0 - 1, c1 - 1) + (O(h0) - O(cb[c0 - 1])) + (O(h1) - O(cb[c1 - 1])));
    cost = std::min(cost, B.zona_part(c0, c1) + (O(cb[c0]) - O(h0)) + (O(cb[c1]) - O(h1)));
    if (c0 > 0)
        cost = std::min(cost, B.zona_part(c0 - 1, c1) + (O(h0) - O(cb[c0 - 1])) + (O(cb[c1]) - O(h1)));
    if (cost > (long)(0.99 * fine)) cost = fine;
    return (double)cost;
}

// Stack depth and short-circuit table of generated programs (the rules of
// fbh_expr_prepare: the right operand of AND / MINUS is skipped on tiles
// where the left one is empty).
void program_tables(const std::vector<int32_t>& prog, const std::vector<uint32_t>& poff, uint32_t nq,
                    std::vector<int32_t>& skip, uint32_t& depth) {
    skip.assign(prog.size(), -1);
    depth = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        uint32_t start[fbk::EXPR_DEPTH];
        int sp = 0;
        for (uint32_t p = poff[q]; p < poff[q + 1]; ++p) {
            const int32_t o = prog[p];
            if (o >= 0) { start[sp++] = p; depth = std::max(depth, (uint32_t)sp); }
            else if (o != fbk::EXPR_NOT) {
                if (o == fbk::EXPR_AND || o == fbk::EXPR_MINUS) skip[start[sp - 1]] = (int32_t)(p + 1);
                --sp;
            }
        }
    }
}

// A batch of range queries on a two-level index as one expression batch:
// each query the program of its option (queries under the scan plan: one
// scan term).  plans_used receives each query's plan.
int two_level_batch(fbh_index* ix, uint32_t nq, const double* lo, const int* lop, const double* hi,
                    const int* rop, const int* plans, int plan, fbh_conj** out, std::vector<int>* plans_used,
                    std::vector<TLProgram>* progs) {
    std::vector<fbk::QueryArgs> h;
    std::vector<int32_t> prog;
    std::vector<uint32_t> toff(nq + 1, 0), poff(nq + 1, 0);
    std::vector<int> used(nq);
    if (progs) progs->assign(nq, TLProgram());
    for (uint32_t i = 0; i < nq; ++i) {
        int p = plans ? plans[i] : plan;
        if (p == FBH_PLAN_AUTO) p = plan_heuristic(ix, lo[i], lop[i], hi[i], rop[i]);
        if (p != FBH_PLAN_SCAN) p = FBH_PLAN_LITERAL;
        used[i] = p;
        TLProgram P;
        if (p == FBH_PLAN_SCAN) {
            fbk::QueryArgs q;
            term_args(fbh_term{ix, lo[i], lop[i], hi[i], rop[i]}, FBH_PLAN_SCAN, q);
            P.terms.push_back(q);
            P.prog.push_back(0);
        } else {
            two_level_program(ix, lo[i], lop[i], hi[i], rop[i], P);
        }
        for (int32_t o : P.prog) prog.push_back(o >= 0 ? (int32_t)(h.size() + (size_t)o) : o);
        for (const auto& q : P.terms) h.push_back(q);
        toff[i + 1] = (uint32_t)h.size();
        poff[i + 1] = (uint32_t)prog.size();
        if (progs) (*progs)[i] = std::move(P);
    }
    std::vector<int32_t> skip;
    uint32_t depth = 0;
    program_tables(prog, poff, nq, skip, depth);
    if (plans_used) *plans_used = used;
    fbh_conj* c = new fbh_conj();
    c->nq = nq; c->nterms = (uint32_t)h.size(); c->ntiles = ix->ntiles; c->groups = ix->groups; c->anb = ix->anb;
    c->depth = depth;
    c->plans = used;                         // the caller's terms are the queries
    c->order.resize(c->nterms);
    for (uint32_t k = 0; k < c->nterms; ++k) c->order[k] = k;
    c->dskip = (int32_t*)dev_alloc(4ull * prog.size());
    int rc = c->dskip ? to_dev(c->dskip, skip.data(), prog.size()) : fail(FBH_ERR_NOMEM, "device allocation failed");
    if (!rc) rc = conj_upload(c, h, toff.data(), prog.data(), poff.data(), (uint32_t)prog.size());
    if (rc) { delete c; return rc; }
    *out = c;
    return FBH_OK;
}


// The coarse section of a fuzz / zona index file, after the relic part
// ending at byte `pos` (src/ixfuzz.cpp:46-143, src/ixzona.cpp:40-100): nc,
// cbounds[nc+1], the coarse bitmaps' positions (offset width w), the
// bitmaps.  Checked on the host like the fine level.
int tl_read(fbh_index* ix, int enc, const uint8_t* file, size_t n, size_t pos, size_t w) {
    auto u32 = [&](size_t at) { uint32_t v; memcpy(&v, file + at, 4); return v; };
    if (pos + 4 > n) return fail(FBH_ERR_ARG, "truncated coarse level");
    const uint32_t nc = u32(pos);
    if (nc == 0 || nc > ix->nobs) return fail(FBH_ERR_ARG, "bad coarse bin count");
    const uint32_t ncb = enc == FBH_ENC_FUZZ ? nc - (nc + 1) / 2 + 1 : nc;
    const size_t cbeg = pos + 4, pbeg = cbeg + 4ull * (nc + 1), pend = pbeg + w * (ncb + 1ull);
    if (pend > n) return fail(FBH_ERR_ARG, "truncated coarse level");
    std::vector<uint32_t> cb(nc + 1);
    for (uint32_t i = 0; i <= nc; ++i) cb[i] = u32(cbeg + 4ull * i);
    if (cb[0] != 0 || cb[nc] != ix->nobs) return fail(FBH_ERR_ARG, "coarse bins do not cover the bitmaps");
    for (uint32_t i = 0; i < nc; ++i)
        if (cb[i + 1] <= cb[i]) return fail(FBH_ERR_ARG, "coarse bins are not ascending");
    std::vector<int64_t> cpos(ncb + 1), woff(ncb + 1);
    for (uint32_t i = 0; i <= ncb; ++i) {
        if (w == 8) memcpy(&cpos[i], file + pbeg + 8ull * i, 8);
        else { int32_t v; memcpy(&v, file + pbeg + 4ull * i, 4); cpos[i] = v; }
    }
    if (cpos[0] < (int64_t)pend || cpos[0] % 4 != 0 || (uint64_t)cpos[ncb] > n)
        return fail(FBH_ERR_ARG, "coarse bitmap positions outside the file");
    for (uint32_t i = 0; i <= ncb; ++i) {
        if (cpos[i] < cpos[0] || (cpos[i] - cpos[0]) % 4 != 0 || (i && cpos[i] < cpos[i - 1]))
            return fail(FBH_ERR_ARG, "bad coarse bitmap position");
        woff[i] = (cpos[i] - cpos[0]) / 4;
    }
    int rc = tl_attach(ix, enc, std::move(cb), cpos, woff, (const uint32_t*)(file + cpos[0]));
    if (rc) return rc;
    ix->toff = tl_fine_offsets(ix);
    return FBH_OK;
}

// Bitmap words a program reads (each term's OR set on its level) plus the
// literal result.
double tl_program_bytes(const fbh_index* ix, const TLProgram& P) {
    double w = 4.0 * (ix->groups + 1);
    for (const fbk::QueryArgs& q : P.terms) {
        const fbh_index* lvl = (ix->coarse && q.tw == ix->coarse->tw) ? ix->coarse : ix;
        auto words = [&](uint32_t b0, uint32_t b1) {
            b1 = std::min(b1, lvl->nobs);
            return b1 > b0 ? lvl->cumser[b1] - lvl->cumser[b0] : 0.0;
        };
        w += words(q.r0, q.r1) + words(q.r2, q.r3);
    }
    return w;
}

double tl_query_bytes(const fbh_index* ix, double lo, int lop, double hi, int rop) {
    TLProgram P;
    two_level_program(ix, lo, lop, hi, rop, P);
    return tl_program_bytes(ix, P);
}

// One range on a two-level index through a one-query expression batch:
// the count, the result bitvector and / or the first `cap` row ids.
int tl_evaluate(fbh_index* ix, double lo, int lop, double hi, int rop, uint64_t* hits, fbh_bitvector** result,
                uint32_t* rids, uint64_t cap) {
    fbh_conj* c = nullptr;
    const int plan = FBH_PLAN_LITERAL;
    int rc = two_level_batch(ix, 1, &lo, &lop, &hi, &rop, nullptr, plan, &c, nullptr, nullptr);
    if (rc) return rc;
    uint64_t off[2] = {0, 0};
    rc = fbh_conj_rids(c, off, rids, rids ? cap : 0);
    if (!rc && hits) *hits = off[1];
    if (!rc && result) rc = fbh_conj_result(c, 0, result);
    delete c;
    return rc;
}

int tl_batch_counts(fbh_index* ix, uint32_t nq, const double* lo, const int* lop, const double* hi,
                    const int* rop, const int* plans, int plan, uint64_t* hits, int* plans_used, float* gpu_ms) {
    fbh_conj* c = nullptr;
    std::vector<int> used;
    int rc = two_level_batch(ix, nq, lo, lop, hi, rop, plans, plan, &c, &used, nullptr);
    if (rc) return rc;
    rc = fbh_conj_run(c, hits, gpu_ms);
    if (!rc && plans_used) for (uint32_t i = 0; i < nq; ++i) plans_used[i] = used[i];
    delete c;
    return rc;
}

int tl_batch_prepare(fbh_index* ix, uint32_t nq, const double* lo, const int* lop, const double* hi,
                     const int* rop, const int* plans, int plan, fbh_conj** out) {
    return two_level_batch(ix, nq, lo, lop, hi, rop, plans, plan, out, nullptr, nullptr);
}

} // namespace
