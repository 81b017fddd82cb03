// C entry points of the index oracle (ctypes).  TEST INFRASTRUCTURE ONLY.
#include "index.hpp"
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace orc;

template <class T>
static T* dup(const std::vector<T>& v) {
    T* p = (T*)std::malloc(sizeof(T) * (v.size() ? v.size() : 1));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

extern "C" {

void* orc_relic_card(const uint32_t* ind, uint32_t n, uint32_t card) { return relic_card(ind, n, card); }
void* orc_relic_values(const double* v, uint32_t n) { return relic_values(v, n); }
void* orc_bin_bounds(const double* v, uint32_t n, const double* b, uint32_t nb) {
    return bin_bounds(v, n, b, nb);
}
void orc_index_free(void* ix) { delete (Index*)ix; }
// the index as one read from a file (offset64 holds serialized positions)
void orc_index_set_offsets_built(void* ix, int v) { ((Index*)ix)->offsets_built = v != 0; }

int orc_index_write(void* ix, double** keys, size_t* nk, int64_t** offs, size_t* no, uint32_t** bms,
                    size_t* nb) {
    std::vector<double> k;
    std::vector<int64_t> o;
    std::vector<word_t> b;
    write_index(*(Index*)ix, k, o, b);
    *keys = dup(k); *nk = k.size();
    *offs = dup(o); *no = o.size();
    *bms = dup(b); *nb = b.size();
    return 0;
}

int orc_index_bounds(void* ix, double** bounds, size_t* n) {
    const Index* p = (Index*)ix;
    *bounds = dup(p->vals);
    *n = p->vals.size();
    return 0;
}

// hits: count and ascending rids
int orc_index_eval(void* ix, const void* col, int type, double lo, int lop, double hi, int rop,
                   uint64_t* count, uint32_t** rids, size_t* nr) {
    try {
        BV r = evaluate(*(Index*)ix, col, type, lo, lop, hi, rop);
        std::vector<word_t> v;
        r.rids(v);
        *count = r.cnt();
        *rids = dup(v);
        *nr = v.size();
        return 0;
    } catch (...) {
        return -2;
    }
}

// estimate: the serialized words of lower and upper (upper cleared: one
// word, 0)
int orc_index_estimate(void* ix, const void* col, int type, double lo, int lop, double hi, int rop,
                       uint32_t** lw, size_t* nl, uint32_t** uw, size_t* nu) {
    try {
        BV lower, upper;
        estimate(*(Index*)ix, col, type, lo, lop, hi, rop, lower, upper);
        std::vector<word_t> a = lower.serial(), b = upper.serial();
        *lw = dup(a); *nl = a.size();
        *uw = dup(b); *nu = b.size();
        return 0;
    } catch (...) {
        return -2;
    }
}

// plan: 0 the reference's choice under `model`, 1 index, 3 scan
int orc_index_eval_plan(void* ix, const void* col, int type, double lo, int lop, double hi, int rop,
                        int plan, int model, uint64_t* count, uint32_t** rids, size_t* nr,
                        int* used) {
    try {
        BV r = evaluate_plan(*(Index*)ix, col, type, lo, lop, hi, rop, plan, model, used);
        std::vector<word_t> v;
        r.rids(v);
        *count = r.cnt();
        *rids = dup(v);
        *nr = v.size();
        return 0;
    } catch (...) {
        return -2;
    }
}

int orc_index_cost(void* ix, const void* col, int type, double lo, int lop, double hi, int rop,
                   int model, double* icost, double* scost, int* plan, int* empty) {
    const Index& x = *(Index*)ix;
    *icost = estimate_cost(x, type, lo, lop, hi, rop, model);
    *scost = scan_cost(x, type, model);
    bool e = false;
    *plan = auto_plan(x, col, type, lo, lop, hi, rop, model, &e);
    *empty = e ? 1 : 0;
    return 0;
}

// --- two-level encodings (fuzz / zona) over a relic
void* orc_two_level_build(void* ix, int enc, uint32_t ncoarse) {
    try {
        return coarsen(*(Index*)ix, enc, ncoarse);
    } catch (...) {
        return nullptr;
    }
}

// The coarse level as a file holds it: fine bitmap positions (bytes, nobs+1),
// cbounds (nc+1), coarse positions (bytes, ncb+1) and the coarse bitmaps'
// serialized words (bitmap i at words[woff[i], woff[i+1])).
void* orc_two_level_file(void* ix, int enc, const int64_t* fine_pos, uint32_t nobs, const uint32_t* cbounds,
                         uint32_t nc, const int64_t* coff, uint32_t ncb, const uint32_t* words, const int64_t* woff) {
    std::vector<int64_t> fp(fine_pos, fine_pos + nobs + 1), co(coff, coff + ncb + 1);
    std::vector<uint32_t> cb(cbounds, cbounds + nc + 1);
    std::vector<std::vector<word_t>> cw(ncb);
    for (uint32_t i = 0; i < ncb; ++i) cw[i].assign(words + woff[i], words + woff[i + 1]);
    return two_level_from_file(*(Index*)ix, enc, fp, cb, co, cw);
}

void orc_two_level_free(void* tl) { delete (TwoLevel*)tl; }

int orc_two_level_info(void* tl, uint32_t** cbounds, size_t* ncb, int64_t** coff, size_t* nco, int64_t** off,
                       size_t* noff) {
    const TwoLevel& t = *(TwoLevel*)tl;
    *cbounds = dup(t.cbounds); *ncb = t.cbounds.size();
    *coff = dup(t.coff); *nco = t.coff.size();
    *off = dup(t.off); *noff = t.off.size();
    return 0;
}

// serialized words of coarse bitmap i
int orc_two_level_cbit(void* tl, uint32_t i, uint32_t** out, size_t* n) {
    const TwoLevel& t = *(TwoLevel*)tl;
    if (i >= t.cbits.size()) return -1;
    std::vector<word_t> v = t.cbits[i]->serial();
    *out = dup(v); *n = v.size();
    return 0;
}

double orc_two_level_cost(void* ix, void* tl, int type, double lo, int lop, double hi, int rop) {
    return two_level_cost(*(Index*)ix, *(TwoLevel*)tl, type, lo, lop, hi, rop);
}

int orc_two_level_eval(void* ix, void* tl, double lo, int lop, double hi, int rop, uint64_t* count,
                       uint32_t** rids, size_t* nr, int* option, uint64_t* loads) {
    try {
        BV r = two_level_eval(*(Index*)ix, *(TwoLevel*)tl, lo, lop, hi, rop, option, loads);
        std::vector<word_t> v;
        r.rids(v);
        *count = r.cnt();
        *rids = dup(v);
        *nr = v.size();
        return 0;
    } catch (...) {
        return -2;
    }
}

} // extern "C"
