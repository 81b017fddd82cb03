/* C++ host binding of libfastbit_hip.so in the shape of FastBit's own
 * classes: fbh::bitvector for ibis::bitvector (src/bitvector.h:62-318),
 * fbh::index for ibis::relic / ibis::bin with the in-memory API around them
 * (src/iapi.h:92-227), fbh::batch for conjunction and expression batches.
 * Header only, over the C ABI in fastbit_hip.h; errors become fbh::error
 * exceptions carrying fbh_last_error() (the reference throws const char*
 * and returns negative codes; no exception crosses the C ABI itself). */
#ifndef FASTBIT_HIP_HPP
#define FASTBIT_HIP_HPP

#include "fastbit_hip.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace fbh {

struct error : std::runtime_error {
    int code;
    error(int rc, const char* what)
        : std::runtime_error(std::string(what) + ": " + (fbh_last_error() ? fbh_last_error() : "")), code(rc) {}
};

inline void check(int rc, const char* what) {
    if (rc != 0) throw error(rc, what);
}

/* ibis::bitvector held in HBM.  The serialized array (write(array_t&),
 * src/bitvector.cpp:2283-2313) is the exchange format both ways. */
class bitvector {
public:
    bitvector() { check(fbh_bv_new(&h_), "fbh_bv_new"); }
    explicit bitvector(const std::vector<uint32_t>& serial) {
        check(fbh_bv_from_serial(serial.data(), (uint32_t)serial.size(), &h_), "fbh_bv_from_serial");
    }
    bitvector(const bitvector& o) { check(fbh_bv_copy(o.h_, &h_), "fbh_bv_copy"); }
    bitvector(bitvector&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    bitvector& operator=(bitvector o) { std::swap(h_, o.h_); return *this; }
    ~bitvector() { if (h_) fbh_bv_free(h_); }

    /* set(val, n), src/bitvector.cpp:256-276 */
    void set(int val, uint32_t n) { check(fbh_bv_set(h_, val, n), "fbh_bv_set"); }
    /* in-place operators, each followed by compress() as the reference's
     * callers do (src/bitvector.cpp:1194-1762) */
    bitvector& operator&=(const bitvector& o) { return op(o, FBH_AND); }
    bitvector& operator|=(const bitvector& o) { return op(o, FBH_OR); }
    bitvector& operator^=(const bitvector& o) { return op(o, FBH_XOR); }
    bitvector& operator-=(const bitvector& o) { return op(o, FBH_MINUS); }
    void compress() { check(fbh_bv_compress(h_), "fbh_bv_compress"); }
    void decompress() { check(fbh_bv_decompress(h_), "fbh_bv_decompress"); }
    void flip() { check(fbh_bv_flip(h_), "fbh_bv_flip"); }
    void adjustSize(uint32_t nv, uint32_t nt) { check(fbh_bv_adjust_size(h_, nv, nt), "fbh_bv_adjust_size"); }
    uint32_t cnt() const { uint32_t c = 0; check(fbh_bv_cnt(h_, &c), "fbh_bv_cnt"); return c; }
    uint32_t size() const { uint32_t n = 0; check(fbh_bv_size(h_, &n), "fbh_bv_size"); return n; }
    /* count(mask), src/bitvector.cpp:1088-1124: ones of *this AND mask */
    uint32_t count(const bitvector& mask) const {
        uint32_t c = 0;
        check(fbh_bv_count(h_, mask.h_, &c), "fbh_bv_count");
        return c;
    }
    /* erase(i, j), src/bitvector.cpp:1012-1081: removes bits [i, j) */
    void erase(uint32_t i, uint32_t j) { check(fbh_bv_erase(h_, i, j), "fbh_bv_erase"); }
    /* the routine the last operator took (FBH_PATH_*) */
    int lastPath() const { return path_; }
    void write(std::vector<uint32_t>& out) const {
        uint32_t n = 0;
        check(fbh_bv_write(h_, nullptr, 0, &n), "fbh_bv_write");
        out.resize(n);
        check(fbh_bv_write(h_, out.data(), n, &n), "fbh_bv_write");
    }
    /* positions of the set bits, ascending (the indexSet walk) */
    std::vector<uint32_t> rids() const {
        uint32_t n = 0;
        check(fbh_bv_rids(h_, nullptr, 0, &n), "fbh_bv_rids");
        std::vector<uint32_t> r(n);
        if (n) check(fbh_bv_rids(h_, r.data(), n, &n), "fbh_bv_rids");
        return r;
    }
    fbh_bitvector* handle() const { return h_; }

private:
    explicit bitvector(fbh_bitvector* h) : h_(h) {}
    bitvector& op(const bitvector& o, int which) {
        check(fbh_bv_op(h_, o.h_, which, &path_), "fbh_bv_op");
        return *this;
    }
    fbh_bitvector* h_ = nullptr;
    int path_ = 0;
    friend bitvector operator&(const bitvector&, const bitvector&);
    friend bitvector operator|(const bitvector&, const bitvector&);
    friend bitvector operator^(const bitvector&, const bitvector&);
    friend bitvector operator-(const bitvector&, const bitvector&);
    friend class batch;
    static bitvector op_new(const bitvector& a, const bitvector& b, int which) {
        fbh_bitvector* r = nullptr;
        check(fbh_bv_op_new(a.h_, b.h_, which, &r, nullptr), "fbh_bv_op_new");
        return bitvector(r);
    }
};

/* out-of-place operators: a new bitvector holding the words the
 * reference's out-of-place dispatch writes, not compressed (the reference
 * returns one the caller deletes, src/bitvector.cpp:1286-1291) */
inline bitvector operator&(const bitvector& a, const bitvector& b) { return bitvector::op_new(a, b, FBH_AND); }
inline bitvector operator|(const bitvector& a, const bitvector& b) { return bitvector::op_new(a, b, FBH_OR); }
inline bitvector operator^(const bitvector& a, const bitvector& b) { return bitvector::op_new(a, b, FBH_XOR); }
inline bitvector operator-(const bitvector& a, const bitvector& b) { return bitvector::op_new(a, b, FBH_MINUS); }

/* In-memory form of an index (fastbit_iapi_deconstruct_index). */
struct index_arrays {
    std::vector<double> keys;
    std::vector<int64_t> offsets;
    std::vector<uint32_t> bitmaps;
};

/* ibis::relic / ibis::bin over one column, with the column, in HBM. */
class index {
public:
    static index relic(const void* col, int type, uint32_t nrows) {
        index ix;
        check(fbh_index_build_relic(col, type, nrows, &ix.h_), "fbh_index_build_relic");
        return ix;
    }
    static index bin(const void* col, int type, uint32_t nrows, const std::vector<double>& bounds) {
        index ix;
        check(fbh_index_build_bin(col, type, nrows, bounds.data(), (uint32_t)bounds.size(), &ix.h_),
              "fbh_index_build_bin");
        return ix;
    }
    static index bin_equal_weight(const void* col, int type, uint32_t nrows, uint32_t nbins) {
        index ix;
        check(fbh_index_build_bin_nbins(col, type, nrows, nbins, &ix.h_), "fbh_index_build_bin_nbins");
        return ix;
    }
    /* ibis::fuzz / ibis::zona: the relic plus a coarse level */
    static index fuzz(const void* col, int type, uint32_t nrows, uint32_t ncoarse = 0) {
        index ix;
        check(fbh_index_build_fuzz(col, type, nrows, ncoarse, &ix.h_), "fbh_index_build_fuzz");
        return ix;
    }
    static index zona(const void* col, int type, uint32_t nrows) {
        index ix;
        check(fbh_index_build_zona(col, type, nrows, &ix.h_), "fbh_index_build_zona");
        return ix;
    }
    /* fastbit_iapi_attach_full_index */
    static index attach(const void* col, int type, uint32_t nrows, const index_arrays& a) {
        index ix;
        check(fbh_index_attach(col, type, nrows, a.keys.data(), a.keys.size(), a.offsets.data(),
                               a.offsets.size(), a.bitmaps.data(), a.bitmaps.size(), &ix.h_),
              "fbh_index_attach");
        return ix;
    }
    /* relic::read / bin::read of an index file */
    static index read(const char* path, const void* col, int type, uint32_t nrows) {
        index ix;
        check(fbh_index_read_file(path, col, type, nrows, &ix.h_), "fbh_index_read_file");
        return ix;
    }
    index(index&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    index& operator=(index&& o) noexcept { std::swap(h_, o.h_); return *this; }
    index(const index&) = delete;
    index& operator=(const index&) = delete;
    ~index() { if (h_) fbh_index_free(h_); }

    /* relic::evaluate / bin::evaluate: the number of hits of
     * lo OP1 x OP2 hi (FBH_OP_* on each side) */
    uint64_t evaluate(double lo, int lop, double hi, int rop, int plan = FBH_PLAN_AUTO) {
        uint64_t hits = 0;
        check(fbh_index_evaluate(h_, lo, lop, hi, rop, plan, &hits, nullptr, nullptr), "fbh_index_evaluate");
        return hits;
    }
    /* the hits as a bitvector */
    bitvector evaluate_bits(double lo, int lop, double hi, int rop, int plan = FBH_PLAN_AUTO) {
        fbh_bitvector* r = nullptr;
        uint64_t hits = 0;
        check(fbh_index_evaluate(h_, lo, lop, hi, rop, plan, &hits, &r, nullptr), "fbh_index_evaluate");
        std::vector<uint32_t> serial;
        uint32_t n = 0;
        int rc = fbh_bv_write(r, nullptr, 0, &n);
        if (!rc) { serial.resize(n); rc = fbh_bv_write(r, serial.data(), n, &n); }
        fbh_bv_free(r);
        check(rc, "fbh_bv_write");
        return bitvector(serial);
    }
    std::vector<uint32_t> rids(double lo, int lop, double hi, int rop, int plan = FBH_PLAN_LITERAL) {
        uint64_t n = 0;
        check(fbh_index_evaluate_rids(h_, lo, lop, hi, rop, plan, nullptr, 0, &n), "fbh_index_evaluate_rids");
        std::vector<uint32_t> r(n);
        if (n) check(fbh_index_evaluate_rids(h_, lo, lop, hi, rop, plan, r.data(), n, &n), "fbh_index_evaluate_rids");
        return r;
    }
    /* fastbit_iapi_get_number_of_hits: ones of sumBins(ib, ie) */
    uint64_t sumBins(uint32_t ib, uint32_t ie) {
        uint64_t n = 0;
        check(fbh_index_sum_bins(h_, ib, ie, &n), "fbh_index_sum_bins");
        return n;
    }
    /* fastbit_iapi_resolve_range: cand0, hit0, hit1, cand1 */
    void locate(double lo, int lop, double hi, int rop, uint32_t& cand0, uint32_t& hit0, uint32_t& hit1,
                uint32_t& cand1) const {
        check(fbh_index_resolve_range(h_, lo, lop, hi, rop, &cand0, &hit0, &hit1, &cand1),
              "fbh_index_resolve_range");
    }
    /* fastbit_iapi_deconstruct_index */
    index_arrays deconstruct() const {
        double* k = nullptr; int64_t* o = nullptr; uint32_t* b = nullptr;
        uint64_t nk = 0, no = 0, nb = 0;
        const int rc = fbh_index_deconstruct(h_, &k, &nk, &o, &no, &b, &nb);
        index_arrays a;
        if (!rc) {
            a.keys.assign(k, k + nk);
            a.offsets.assign(o, o + no);
            a.bitmaps.assign(b, b + nb);
        }
        fbh_free(k); fbh_free(o); fbh_free(b);
        check(rc, "fbh_index_deconstruct");
        return a;
    }
    uint32_t numBitvectors() const { uint32_t n = 0; check(fbh_index_info(h_, &n, nullptr, nullptr), "fbh_index_info"); return n; }
    uint32_t getNRows() const { uint32_t n = 0; check(fbh_index_info(h_, nullptr, &n, nullptr), "fbh_index_info"); return n; }
    const fbh_index* handle() const { return h_; }

private:
    index() = default;
    fbh_index* h_ = nullptr;
};

/* A prepared batch of conjunctions (fbh_conj_prepare) or boolean
 * expressions (fbh_expr_prepare), replayed from a hipGraph. */
class batch {
public:
    static batch conjunctions(const std::vector<fbh_term>& terms, const std::vector<uint32_t>& qoff,
                              int plan = FBH_PLAN_AUTO) {
        batch b;
        check(fbh_conj_prepare(terms.data(), qoff.data(), (uint32_t)qoff.size() - 1, plan, &b.h_),
              "fbh_conj_prepare");
        b.nq_ = (uint32_t)qoff.size() - 1;
        return b;
    }
    static batch expressions(const std::vector<fbh_term>& terms, const std::vector<uint32_t>& toff,
                             const std::vector<int32_t>& prog, const std::vector<uint32_t>& poff,
                             int plan = FBH_PLAN_AUTO) {
        batch b;
        check(fbh_expr_prepare(terms.data(), toff.data(), prog.data(), poff.data(), (uint32_t)poff.size() - 1,
                               plan, &b.h_),
              "fbh_expr_prepare");
        b.nq_ = (uint32_t)poff.size() - 1;
        return b;
    }
    batch(batch&& o) noexcept : h_(o.h_), nq_(o.nq_) { o.h_ = nullptr; }
    batch(const batch&) = delete;
    batch& operator=(const batch&) = delete;
    ~batch() { if (h_) fbh_conj_free(h_); }
    /* hit counts of every query */
    std::vector<uint64_t> run(float* gpu_ms = nullptr) {
        std::vector<uint64_t> h(nq_);
        check(fbh_conj_run(h_, h.data(), gpu_ms), "fbh_conj_run");
        return h;
    }
    void tune() {
        int t = 0; uint32_t b = 0, tpi = 0; float ms = 0;
        check(fbh_conj_tune(h_, &t, &b, &tpi, &ms), "fbh_conj_tune");
    }
    /* row ids of every query (countQuery::getHitRows per query): query q's
     * ids are rids[offsets[q] .. offsets[q+1]) */
    void hit_rows(std::vector<uint64_t>& offsets, std::vector<uint32_t>& rids) {
        offsets.assign(nq_ + 1, 0);
        check(fbh_conj_rids(h_, offsets.data(), nullptr, 0), "fbh_conj_rids");
        rids.resize(offsets[nq_]);
        check(fbh_conj_rids(h_, offsets.data(), rids.data(), rids.size()), "fbh_conj_rids");
    }
    /* the hit vector of query q (countQuery::getHitVector) */
    bitvector hit_vector(uint32_t q) {
        fbh_bitvector* v = nullptr;
        check(fbh_conj_result(h_, q, &v), "fbh_conj_result");
        return bitvector(v);
    }

private:
    batch() = default;
    fbh_conj* h_ = nullptr;
    uint32_t nq_ = 0;
};

} // namespace fbh

#endif
