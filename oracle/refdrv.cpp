// Thin C driver over the real FastBit library (built from /root/reference by
// oracle/ref.mk into oracle/_ref/).  TEST INFRASTRUCTURE ONLY: it pins the
// oracle restatement and writes the golden fixtures; nothing in the product
// links it.  Every function calls the reference's own public C++ API.
#include "ibis.h"
#include "irelic.h"
#include "ibin.h"
#include "iapi.h"
#include <chrono>
#include <cstdio>
#include <string>
#include <cstdlib>
#include <cstring>
#include <vector>

static bool g_inited = false;
static void ensure_init() {
    if (!g_inited) {
        ibis::gVerbose = -1;
        ibis::init();
        g_inited = true;
    }
}

static int hand_out(const uint32_t* p, size_t n, uint32_t** out, size_t* nout) {
    *nout = n;
    *out = (uint32_t*)std::malloc(sizeof(uint32_t) * (n ? n : 1));
    if (!*out) return -1;
    if (n) std::memcpy(*out, p, sizeof(uint32_t) * n);
    return 0;
}

// Deep copy of a serialized bitvector into a reference bitvector.
static void load(ibis::bitvector& bv, const uint32_t* a, size_t n) {
    ibis::array_t<ibis::bitvector::word_t> arr(n);
    if (n) std::memcpy(arr.begin(), a, n * sizeof(uint32_t));
    ibis::bitvector tmp(arr);
    bv.copy(tmp);
}

extern "C" {

void ref_free(void* p) { std::free(p); }

int ref_wah_op(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
               int op, int compress, uint32_t** out, size_t* nout) {
    ensure_init();
    try {
        ibis::bitvector x, y;
        load(x, a, na);
        load(y, b, nb);
        switch (op) {
        case 0: x &= y; break;
        case 1: x |= y; break;
        case 2: x ^= y; break;
        default: x -= y; break;
        }
        if (compress) x.compress();
        ibis::array_t<ibis::bitvector::word_t> arr;
        x.write(arr);
        return hand_out(arr.begin(), arr.size(), out, nout);
    } catch (...) {
        return -2;
    }
}

// The out-of-place operators a & b, a | b, a ^ b, a - b
// (src/bitvector.cpp:1292, 1488, 1624, 1768): a new bitvector, not
// compressed.
int ref_wah_op_new(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
                   int op, uint32_t** out, size_t* nout) {
    ensure_init();
    try {
        ibis::bitvector x, y;
        load(x, a, na);
        load(y, b, nb);
        ibis::bitvector* r = nullptr;
        switch (op) {
        case 0: r = x & y; break;
        case 1: r = x | y; break;
        case 2: r = x ^ y; break;
        default: r = x - y; break;
        }
        ibis::array_t<ibis::bitvector::word_t> arr;
        r->write(arr);
        delete r;
        return hand_out(arr.begin(), arr.size(), out, nout);
    } catch (...) {
        return -2;
    }
}

// bitvector::count(mask) (src/bitvector.cpp:1088-1124)
int ref_wah_count(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* cnt) {
    ensure_init();
    try {
        ibis::bitvector x, y;
        load(x, a, na);
        load(y, b, nb);
        (void) x.size(); (void) y.size();
        *cnt = x.count(y);
        return 0;
    } catch (...) {
        return -2;
    }
}

// bitvector::erase(i, j) (src/bitvector.cpp:1012-1081) of a vector whose size
// has been counted (erase tests j against nbits as it stands).
int ref_wah_erase(const uint32_t* a, size_t na, uint32_t i, uint32_t j, uint32_t** out, size_t* nout) {
    ensure_init();
    try {
        ibis::bitvector x;
        load(x, a, na);
        (void) x.size();       // erase reads nbits without counting it
        x.erase(i, j);
        ibis::array_t<ibis::bitvector::word_t> arr;
        x.write(arr);
        return hand_out(arr.begin(), arr.size(), out, nout);
    } catch (...) {
        return -2;
    }
}

// The in-place operator followed by compress() on copies of a, `reps`
// times: the mean milliseconds of the op and compress alone (the copy of a
// is made before the clock starts), and the result's word count.
int ref_wah_time_op(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, int op, int reps,
                    double* ms, uint32_t* nwords) {
    ensure_init();
    try {
        ibis::bitvector x0, y;
        load(x0, a, na);
        load(y, b, nb);
        (void) x0.size(); (void) y.size();
        double tot = 0.0;
        for (int r = 0; r < reps; ++r) {
            ibis::bitvector x;
            x.copy(x0);
            const auto t0 = std::chrono::steady_clock::now();
            switch (op) {
            case 0: x &= y; break;
            case 1: x |= y; break;
            case 2: x ^= y; break;
            default: x -= y; break;
            }
            x.compress();
            tot += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (r == 0 && nwords) {
                ibis::array_t<ibis::bitvector::word_t> arr;
                x.write(arr);
                *nwords = (uint32_t)arr.size();
            }
        }
        *ms = tot / (reps > 0 ? reps : 1);
        return 0;
    } catch (...) {
        return -2;
    }
}

int ref_wah_from_bits(const uint8_t* bits, size_t n, uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    for (size_t i = 0; i < n; ++i) x += (bits[i] ? 1 : 0);
    ibis::array_t<ibis::bitvector::word_t> arr;
    x.write(arr);
    return hand_out(arr.begin(), arr.size(), out, nout);
}

int ref_wah_cnt(const uint32_t* a, size_t na, uint32_t* cnt, uint32_t* size,
                uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    load(x, a, na);
    *size = x.size();
    *cnt = x.cnt();
    ibis::array_t<ibis::bitvector::word_t> arr;
    x.write(arr);
    return hand_out(arr.begin(), arr.size(), out, nout);
}

int ref_wah_compress(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    load(x, a, na);
    x.compress();
    ibis::array_t<ibis::bitvector::word_t> arr;
    x.write(arr);
    return hand_out(arr.begin(), arr.size(), out, nout);
}

int ref_wah_decompress(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    load(x, a, na);
    x.decompress();
    ibis::array_t<ibis::bitvector::word_t> arr;
    x.write(arr);
    return hand_out(arr.begin(), arr.size(), out, nout);
}

int ref_wah_flip(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    load(x, a, na);
    (void) x.size();
    x.flip();
    ibis::array_t<ibis::bitvector::word_t> arr;
    x.write(arr);
    return hand_out(arr.begin(), arr.size(), out, nout);
}

int ref_wah_rids(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    load(x, a, na);
    std::vector<uint32_t> r;
    for (ibis::bitvector::indexSet is = x.firstIndexSet(); is.nIndices() > 0; ++is) {
        const ibis::bitvector::word_t* ii = is.indices();
        if (is.isRange()) {
            for (ibis::bitvector::word_t k = ii[0]; k < ii[1]; ++k) r.push_back(k);
        } else {
            for (unsigned k = 0; k < is.nIndices(); ++k) r.push_back(ii[k]);
        }
    }
    return hand_out(r.data(), r.size(), out, nout);
}

} // extern "C"

// ---------------------------------------------------------------------------
// Indexes through the reference's own API.
template <class T>
static T* dupv(const T* p, size_t n) {
    T* q = (T*)std::malloc(sizeof(T) * (n ? n : 1));
    if (q && n) std::memcpy(q, p, sizeof(T) * n);
    return q;
}

extern "C" {

// relic(col, card, ind) (src/irelic.cpp:160-211), then relic::write
int ref_relic_card(const uint32_t* ind, size_t n, uint32_t card, double** keys, size_t* nk,
                   int64_t** offs, size_t* no, uint32_t** bms, size_t* nb) {
    ensure_init();
    try {
        ibis::array_t<uint32_t> arr(n);
        if (n) std::memcpy(arr.begin(), ind, 4 * n);
        ibis::relic rl(0, card, arr);
        ibis::array_t<double> k;
        ibis::array_t<int64_t> o;
        ibis::array_t<uint32_t> b;
        rl.write(k, o, b);
        *keys = dupv(k.begin(), k.size()); *nk = k.size();
        *offs = dupv(o.begin(), o.size()); *no = o.size();
        *bms = dupv(b.begin(), b.size()); *nb = b.size();
        return 0;
    } catch (...) {
        return -2;
    }
}

// fastbit_iapi_register_array + fastbit_iapi_build_index +
// fastbit_iapi_deconstruct_index (src/iapi.cpp:1893-1953).  dtype is a
// FastBitDataType.  The array stays registered under `name` for queries.
int ref_iapi_index(const char* name, int dtype, void* data, uint64_t n, const char* spec,
                   double** keys, uint64_t* nk, int64_t** offs, uint64_t* no, uint32_t** bms,
                   uint64_t* nb) {
    ensure_init();
    int ierr = fastbit_iapi_register_array(name, (FastBitDataType)dtype, data, n);
    if (ierr < 0) return -10 + ierr;
    ierr = fastbit_iapi_build_index(name, spec);
    if (ierr < 0) return -20 + ierr;
    return fastbit_iapi_deconstruct_index(name, keys, nk, offs, no, bms, nb);
}

// lo OP1 x OP2 hi through selections (src/iapi.cpp:1265-1562, 2062-2090);
// ops: 0 none, 1 <, 2 <=.  Returns the hit count; rids (ascending) into buf.
int64_t ref_iapi_query(const char* name, double lo, int lop, double hi, int rop, uint64_t* buf,
                       uint64_t cap) {
    ensure_init();
    FastBitSelectionHandle a = 0, b = 0, h = 0;
    if (lop == 1) a = fastbit_selection_osr(name, FastBitCompareGreater, lo);
    else if (lop == 2) a = fastbit_selection_osr(name, FastBitCompareGreaterEqual, lo);
    if (rop == 1) b = fastbit_selection_osr(name, FastBitCompareLess, hi);
    else if (rop == 2) b = fastbit_selection_osr(name, FastBitCompareLessEqual, hi);
    if (a && b) h = fastbit_selection_combine(a, FastBitCombineAnd, b);
    else h = a ? a : b;
    if (!h) return -1;
    int64_t cnt = fastbit_selection_evaluate(h);
    if (cnt > 0 && buf) {
        int64_t got = fastbit_selection_get_coordinates(h, buf, cap, 0);
        if (got < 0) cnt = got;
    }
    fastbit_selection_free(h);
    return cnt;
}

void ref_iapi_free_all() { fastbit_iapi_free_all(); }

} // extern "C"

extern "C" {

// fastbit_iapi_register_array + fastbit_iapi_attach_full_index
// (src/iapi.cpp:2037-2060): the reference queries an index built elsewhere
// (no copy: the arrays must outlive the registration).
int ref_iapi_attach(const char* name, int dtype, void* data, uint64_t n, double* keys, uint64_t nk,
                    int64_t* offs, uint64_t no, uint32_t* bms, uint64_t nb) {
    ensure_init();
    int ierr = fastbit_iapi_register_array(name, (FastBitDataType)dtype, data, n);
    if (ierr < 0) return -10 + ierr;
    return fastbit_iapi_attach_full_index(name, keys, nk, offs, no, bms, nb);
}

} // extern "C"

extern "C" {

// One two-sided range lo OP1 x OP2 hi as a single qContinuousRange term
// (the form a where clause "lo <= x < hi" takes), evaluated through
// fastbit_selection_evaluate (src/iapi.cpp:1383-1418).  ops: 0 none, 1 <,
// 2 <=.  Returns the hit count; rids (ascending) into buf.
int64_t ref_iapi_query_range(const char* name, double lo, int lop, double hi, int rop, uint64_t* buf,
                             uint64_t cap) {
    ensure_init();
    auto op = [](int o) {
        return o == 1 ? ibis::qExpr::OP_LT : o == 2 ? ibis::qExpr::OP_LE : ibis::qExpr::OP_UNDEFINED;
    };
    ibis::qContinuousRange* q = new ibis::qContinuousRange(lo, op(lop), name, op(rop), hi);
    FastBitSelectionHandle h = static_cast<FastBitSelectionHandle>(q);
    int64_t cnt = fastbit_selection_evaluate(h);
    if (cnt > 0 && buf) {
        int64_t got = fastbit_selection_get_coordinates(h, buf, cap, 0);
        if (got < 0) cnt = got;
    }
    fastbit_selection_free(h);
    return cnt;
}

} // extern "C"

extern "C" void ref_set_verbose(int v) { ensure_init(); ibis::gVerbose = v; }

// The boundaries of the bin index built for a registered array
// (ibis::bin::binBoundaries through the column's index lock).
#include "bord.h"
ibis::bord::column* __fastbit_iapi_array_by_name(const char* name);

// index::estimate(expr, lower, upper) of the index of a registered array
// (bin::estimate, src/ibin.cpp:6722-6829; relic: src/irelic.h:58-62): the
// serialized words of lower and upper.
extern "C" int ref_iapi_estimate(const char* name, double lo, int lop, double hi, int rop, uint32_t** lw,
                                 size_t* nl, uint32_t** uw, size_t* nu) {
    ensure_init();
    ibis::bord::column* col = __fastbit_iapi_array_by_name(name);
    if (col == 0) return -1;
    ibis::column::indexLock lock(col, "ref_iapi_estimate");
    const ibis::index* ix = lock.getIndex();
    if (ix == 0) return -2;
    auto op = [](int o) {
        return o == 1 ? ibis::qExpr::OP_LT : o == 2 ? ibis::qExpr::OP_LE : ibis::qExpr::OP_UNDEFINED;
    };
    try {
        ibis::qContinuousRange q(lo, op(lop), name, op(rop), hi);
        ibis::bitvector lower, upper;
        ix->estimate(q, lower, upper);
        ibis::array_t<ibis::bitvector::word_t> a, b;
        lower.write(a);
        upper.write(b);
        if (hand_out(a.begin(), a.size(), lw, nl)) return -3;
        return hand_out(b.begin(), b.size(), uw, nu);
    } catch (...) {
        return -4;
    }
}

extern "C" int ref_iapi_bin_bounds(const char* name, double** out, uint64_t* n) {
    ensure_init();
    ibis::bord::column* col = __fastbit_iapi_array_by_name(name);
    if (col == 0) return -1;
    ibis::column::indexLock lock(col, "ref_iapi_bin_bounds");
    const ibis::index* ix = lock.getIndex();
    if (ix == 0) return -2;
    std::vector<double> b;
    ix->binBoundaries(b);
    *n = b.size();
    *out = dupv(b.data(), b.size());
    return 0;
}

// ---------------------------------------------------------------------------
// Index files.  An index with no column attached writes to a path that ends
// in ".idx" as given (index::indexFileName, src/index.cpp:1481-1487), so the
// index built for a registered array is copied into an object with col == 0
// and written from there.
namespace {
struct FileRelic : ibis::relic {
    explicit FileRelic(const ibis::relic& r) : ibis::relic(r) { col = 0; }
};
struct FileBin : ibis::bin {
    explicit FileBin(const ibis::bin& b) : ibis::bin(b) { col = 0; }
};

// An index read from a file whose bitmaps are loaded as the file holds
// them: a zero-length range stays a null bitmap (index::activate would make
// it an all-zero one, src/index.cpp:4984-4987).
template <class B>
struct Opened : B {
    explicit Opened(const char* path) : B(0, path) {}
    void load_present() {
        const size_t n = this->bits.size();
        for (size_t i = 0; i < n; ++i) {
            int64_t a = 0, e = 0;
            if (this->offset64.size() > n) { a = this->offset64[i]; e = this->offset64[i + 1]; }
            else if (this->offset32.size() > n) { a = this->offset32[i]; e = this->offset32[i + 1]; }
            if (e > a) this->activate((uint32_t)i);
        }
    }
};

ibis::qExpr::COMPARE cmp_op(int o) {
    return o == 1 ? ibis::qExpr::OP_LT : o == 2 ? ibis::qExpr::OP_LE : ibis::qExpr::OP_UNDEFINED;
}

// The index file at path, read by relic::read or bin::read by its header.
ibis::index* read_file(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return 0;
    char h[8];
    const size_t got = std::fread(h, 1, 8, f);
    std::fclose(f);
    if (got != 8) return 0;
    if (h[5] == (char)ibis::index::RELIC) return new ibis::relic(0, path);
    if (h[5] == (char)ibis::index::BINNING) return new ibis::bin(0, path);
    return 0;
}
} // namespace

extern "C" {

// relic::write / bin::write (src/irelic.cpp:354-456, src/ibin.cpp:5530-5654)
// of the index built for the registered array `name`.
int ref_iapi_write_index_file(const char* name, const char* path) {
    ensure_init();
    ibis::bord::column* c = __fastbit_iapi_array_by_name(name);
    if (c == 0) return -1;
    ibis::column::indexLock lock(c, "ref_iapi_write_index_file");
    const ibis::index* ix = lock.getIndex();
    if (ix == 0) return -2;
    try {
        if (ix->type() == ibis::index::RELIC) {
            FileRelic r(*static_cast<const ibis::relic*>(ix));
            return r.write(path);
        }
        if (ix->type() == ibis::index::BINNING) {
            FileBin b(*static_cast<const ibis::bin*>(ix));
            return b.write(path);
        }
    } catch (...) {
        return -4;
    }
    return -3;
}

// relic(0, card, ind) (src/irelic.cpp:160-211) written to an index file.
int ref_relic_card_file(const uint32_t* ind, size_t n, uint32_t card, const char* path) {
    ensure_init();
    try {
        ibis::array_t<uint32_t> arr(n);
        if (n) std::memcpy(arr.begin(), ind, 4 * n);
        ibis::relic rl(0, card, arr);
        return rl.write(path);
    } catch (...) {
        return -2;
    }
}

// The in-memory form (relic::write / bin::write into arrays,
// src/irelic.cpp:644-667, src/ibin.cpp:5836-5860) of an index file read by
// the reference; *kind is 0 for bin, 1 for relic.
int ref_file_triple(const char* path, int* kind, double** keys, uint64_t* nk, int64_t** offs,
                    uint64_t* no, uint32_t** bms, uint64_t* nb) {
    ensure_init();
    FILE* f = std::fopen(path, "rb");
    if (!f) return -1;
    char h[8];
    const size_t got = std::fread(h, 1, 8, f);
    std::fclose(f);
    if (got != 8) return -1;
    ibis::array_t<double> k;
    ibis::array_t<int64_t> o;
    ibis::array_t<uint32_t> b;
    int rc = -1;
    try {
        if (h[5] == (char)ibis::index::RELIC) {
            Opened<ibis::relic> ix(path);
            ix.load_present();
            rc = ix.write(k, o, b);
            *kind = 1;
        } else if (h[5] == (char)ibis::index::BINNING) {
            Opened<ibis::bin> ix(path);
            ix.load_present();
            rc = ix.write(k, o, b);
            *kind = 0;
        }
    } catch (...) {
        rc = -2;
    }
    if (rc < 0) return rc;
    *keys = dupv(k.begin(), k.size()); *nk = k.size();
    *offs = dupv(o.begin(), o.size()); *no = o.size();
    *bms = dupv(b.begin(), b.size()); *nb = b.size();
    return 0;
}

// relic::locate / bin::locate of lo OP1 x OP2 hi on an index file read by
// the reference: out = cand0, hit0, hit1, cand1 (fastbit_iapi_resolve_range's
// order, src/iapi.cpp:1981-2009).
int ref_file_locate(const char* path, double lo, int lop, double hi, int rop, uint32_t* out) {
    ensure_init();
    ibis::index* ix = read_file(path);
    if (ix == 0) return -1;
    ibis::qContinuousRange cr(lo, cmp_op(lop), "_", cmp_op(rop), hi);
    uint32_t c0 = 0, h0 = 0, h1 = 0, c1 = 0;
    if (ix->type() == ibis::index::RELIC) {
        static_cast<const ibis::relic*>(ix)->locate(cr, h0, h1);
        c0 = h0; c1 = h1;
    } else {
        static_cast<const ibis::bin*>(ix)->locate(cr, c0, c1, h0, h1);
    }
    out[0] = c0; out[1] = h0; out[2] = h1; out[3] = c1;
    delete ix;
    return 0;
}

// index::sumBins(ib, ie) (src/index.cpp:6450-7019) of an index file read by
// the reference: the number of ones.
int64_t ref_file_sum_bins(const char* path, uint32_t ib, uint32_t ie) {
    ensure_init();
    ibis::index* ix = read_file(path);
    if (ix == 0) return -1;
    int64_t cnt = -2;
    try {
        ibis::bitvector res;
        ix->sumBins(ib, ie, res);
        cnt = res.cnt();
    } catch (...) {
    }
    delete ix;
    return cnt;
}

} // extern "C"

extern "C" {

// A boolean expression over range terms on registered arrays, built with
// fastbit_selection_combine (src/iapi.cpp:1285-1339) from a postfix program
// (item >= 0: term; -1 AND, -2 OR, -3 XOR, -4 NAND, -5 NOR, -6 NOT,
// -7 MINUS) and counted by fastbit_selection_evaluate (countQuery).  Each
// term is one qContinuousRange lo OP1 names[i] OP2 hi.  Returns the hit
// count, or a negative code.
int64_t ref_iapi_expr(const char* const* names, const double* lo, const int* lop, const double* hi,
                      const int* rop, const int32_t* prog, uint32_t nprog) {
    ensure_init();
    std::vector<ibis::qExpr*> st;
    for (uint32_t p = 0; p < nprog; ++p) {
        const int32_t op = prog[p];
        if (op >= 0) {
            st.push_back(new ibis::qContinuousRange(lo[op], cmp_op(lop[op]), names[op], cmp_op(rop[op]), hi[op]));
            continue;
        }
        if (op == -6) {
            if (st.empty()) return -1;
            ibis::qExpr* n = new ibis::qExpr(ibis::qExpr::LOGICAL_NOT);
            n->setLeft(st.back());
            st.back() = n;
            continue;
        }
        if (st.size() < 2) return -1;
        ibis::qExpr* b = st.back(); st.pop_back();
        ibis::qExpr* a = st.back(); st.pop_back();
        ibis::qExpr* r = 0;
        if (op == -7) {
            r = new ibis::qExpr(ibis::qExpr::LOGICAL_MINUS);
            r->setLeft(a);
            r->setRight(b);
        } else {
            const FastBitCombineType t = op == -1 ? FastBitCombineAnd : op == -2 ? FastBitCombineOr
                                         : op == -3 ? FastBitCombineXor : op == -4 ? FastBitCombineNand
                                         : FastBitCombineNor;
            r = static_cast<ibis::qExpr*>(fastbit_selection_combine(a, t, b));
        }
        if (r == 0) return -2;
        st.push_back(r);
    }
    if (st.size() != 1) return -1;
    FastBitSelectionHandle h = static_cast<FastBitSelectionHandle>(st[0]);
    const int64_t cnt = fastbit_selection_evaluate(h);
    fastbit_selection_free(h);
    return cnt;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Two-level encodings over a relic: ibis::fuzz (interval-equality) and
// ibis::zona (equality-equality), src/irelic.h:540-712.
namespace {

// ibis::index::col is protected; a pointer to it taken inside a derived
// class reaches it on any index.
struct IndexCol : ibis::index {
    static const ibis::column* ibis::index::*member() { return &IndexCol::col; }
};

// An index file through the file manager (fileManager::getFile, then the
// storage constructors fuzz(c, st) / zona(c, st) / relic(c, st) as
// index::create uses them, src/index.cpp:781-795): the bitmaps are copied
// out of the mapped file when an evaluation first needs them.  (Read with
// fuzz::read(name) instead, which loads bitmaps by file name, the fuzz_i32
// fixture's file gives 521 hits for 283.48 <= x <= 313.36, where 1,233 rows
// qualify; through the file manager every term matches the column.)
ibis::relic* read_two_level(const char* path) {
    ibis::fileManager::storage* st = 0;
    if (ibis::fileManager::instance().getFile(path, &st) != 0 || st == 0 || st->size() < 8) return 0;
    const char t = st->begin()[5];
    if (t == (char)ibis::index::FUZZ) return new ibis::fuzz(0, st);
    if (t == (char)ibis::index::ZONA) return new ibis::zona(0, st);
    if (t == (char)ibis::index::RELIC) return new ibis::relic(0, st);
    return 0;
}

// fuzz::cbits / zona::cbits are private; an explicit instantiation may name
// a private member, and the friend function it defines hands out the
// pointer to it.
template <class Tag, typename Tag::type M>
struct Rob {
    friend typename Tag::type get(Tag) { return M; }
};
struct FuzzCbits {
    typedef ibis::array_t<ibis::bitvector*> ibis::fuzz::*type;
    friend type get(FuzzCbits);
};
struct ZonaCbits {
    typedef ibis::array_t<ibis::bitvector*> ibis::zona::*type;
    friend type get(ZonaCbits);
};
template struct Rob<FuzzCbits, &ibis::fuzz::cbits>;
template struct Rob<ZonaCbits, &ibis::zona::cbits>;

// Bit i set when coarse bitmap i is in memory.
uint64_t coarse_loaded(const ibis::relic* r) {
    const ibis::array_t<ibis::bitvector*>* cb = 0;
    if (r->type() == ibis::index::FUZZ) cb = &(static_cast<const ibis::fuzz*>(r)->*get(FuzzCbits()));
    else if (r->type() == ibis::index::ZONA) cb = &(static_cast<const ibis::zona*>(r)->*get(ZonaCbits()));
    uint64_t m = 0;
    if (cb)
        for (size_t i = 0; i < cb->size() && i < 64; ++i)
            if ((*cb)[i] != 0) m |= 1ull << i;
    return m;
}

} // namespace

extern "C" {

// fuzz::write / zona::write (src/ixfuzz.cpp:1098-1209, src/ixzona.cpp:
// 1060-1130) of the index built for the registered array `name` (built with
// an "<encoding interval-equality/>" or "equality-equality" spec).  The
// index is written as one with no column attached, so that the path is used
// as given.
int ref_iapi_write_coarse_file(const char* name, const char* path) {
    ensure_init();
    ibis::bord::column* c = __fastbit_iapi_array_by_name(name);
    if (c == 0) return -1;
    ibis::column::indexLock lock(c, "ref_iapi_write_coarse_file");
    ibis::index* ix = const_cast<ibis::index*>(lock.getIndex());
    if (ix == 0) return -2;
    if (ix->type() != ibis::index::FUZZ && ix->type() != ibis::index::ZONA) return -3;
    const ibis::column* saved = ix->*IndexCol::member();
    ix->*IndexCol::member() = 0;
    int rc = -4;
    try {
        rc = ix->write(path);
    } catch (...) {
    }
    ix->*IndexCol::member() = saved;
    return rc;
}

// fuzz::evaluate / zona::evaluate and estimateCost (src/ixfuzz.cpp:767-1095,
// src/ixzona.cpp:580-1040) of lo OP1 x OP2 hi, on the index built for the
// registered array `name` (path null) or on the index file at `path` (the
// array `name` is attached as its column, for the null mask).  An index
// read from a file loads its bitmaps when an evaluation first needs them:
// *loaded gets bit i set for every coarse bitmap evaluate read, which names
// the option it took.  Returns the hit count.
int64_t ref_two_level_eval(const char* name, const char* path, double lo, int lop, double hi, int rop,
                           double* cost, uint64_t* loaded) {
    ensure_init();
    ibis::bord::column* c = __fastbit_iapi_array_by_name(name);
    if (c == 0) return -1;
    ibis::qContinuousRange cr(lo, cmp_op(lop), name, cmp_op(rop), hi);
    int64_t cnt = -2;
    *loaded = 0;
    ibis::relic* owned = 0;
    const ibis::relic* ix = 0;
    ibis::column::indexLock* lock = 0;
    if (path) {
        owned = read_two_level(path);
        if (owned == 0) return -3;
        owned->*IndexCol::member() = c;
        ix = owned;
    } else {
        lock = new ibis::column::indexLock(c, "ref_two_level_eval");
        ix = static_cast<const ibis::relic*>(lock->getIndex());
        if (ix == 0) { delete lock; return -4; }
    }
    try {
        *cost = ix->estimateCost(cr);
        ibis::bitvector res;
        cnt = ix->evaluate(cr, res);
        if (path) *loaded = coarse_loaded(ix);
    } catch (...) {
        cnt = -5;
    }
    if (owned) { owned->*IndexCol::member() = 0; delete owned; }
    delete lock;
    return cnt;
}

} // extern "C"
