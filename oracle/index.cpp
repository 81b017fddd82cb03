// CPU oracle: relic / bin restatement.  TEST INFRASTRUCTURE ONLY.
#include "index.hpp"
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <map>

namespace orc {

// src/irelic.cpp:160-211
Index* relic_card(const uint32_t* ind, uint32_t n, uint32_t card) {
    Index* ix = new Index();
    ix->kind = 1;
    ix->nrows = n;
    ix->vals.resize(card);
    ix->bits.resize(card);
    for (uint32_t i = 0; i < card; ++i) { ix->vals[i] = i; ix->bits[i] = new BV(); }
    for (uint32_t i = 0; i < n; ++i)
        if (ind[i] < card) ix->bits[ind[i]]->setBit(i, 1);
    for (uint32_t i = 0; i < card; ++i) ix->bits[i]->adjustSize(0, n);
    return ix;
}

// src/index.cpp:2552-2611 (a std::map from value to bitvector, setBit per
// row), then every bitmap padded to nrows
Index* relic_values(const double* v, uint32_t n) {
    std::map<double, BV*> m;
    for (uint32_t i = 0; i < n; ++i) {
        auto it = m.find(v[i]);
        BV* b;
        if (it == m.end()) { b = new BV(); m[v[i]] = b; } else b = it->second;
        b->setBit(i, 1);
    }
    Index* ix = new Index();
    ix->kind = 1;
    ix->nrows = n;
    for (auto& kv : m) {
        kv.second->adjustSize(0, n);
        ix->vals.push_back(kv.first);
        ix->bits.push_back(kv.second);
    }
    return ix;
}

// src/ibin.cpp:673-738
static uint32_t bin_locate(const std::vector<double>& bounds, double val) {
    const uint32_t nobs = (uint32_t)bounds.size();
    if (nobs == 0) return 0;
    if (!(val >= bounds[0])) return 0;
    if (!(val < bounds[nobs - 1])) return bounds[nobs - 1] < DBL_MAX ? nobs : nobs - 1;
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

// src/ibin.cpp:3350-3449
Index* bin_bounds(const double* v, uint32_t n, const double* bounds, uint32_t nb) {
    Index* ix = new Index();
    ix->kind = 2;
    ix->nrows = n;
    ix->vals.assign(bounds, bounds + nb);
    uint32_t nobs = nb;
    ix->bits.resize(nobs);
    ix->minval.assign(nobs, DBL_MAX);
    ix->maxval.assign(nobs, -DBL_MAX);
    for (uint32_t i = 0; i < nobs; ++i) ix->bits[i] = new BV();
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t j = bin_locate(ix->vals, v[i]);
        if (j < nobs) {
            ix->bits[j]->setBit(i, 1);
            if (ix->minval[j] > v[i]) ix->minval[j] = v[i];
            if (ix->maxval[j] < v[i]) ix->maxval[j] = v[i];
        }
    }
    for (uint32_t i = 0; i < nobs; ++i) {
        if (ix->bits[i]->cnt() > 0) ix->bits[i]->adjustSize(0, n);
        else { delete ix->bits[i]; ix->bits[i] = nullptr; }
    }
    if (nobs > 0) {
        --nobs;
        uint32_t k = 1;
        for (uint32_t i = 1; i < nobs; ++i) {
            if (ix->bits[i] != nullptr) {
                if (i > k) {
                    ix->vals[k] = ix->vals[i];
                    ix->minval[k] = ix->minval[i];
                    ix->maxval[k] = ix->maxval[i];
                    ix->bits[k] = ix->bits[i];
                }
                ++k;
            }
        }
        if (nobs > k) {
            ix->vals[k] = ix->vals[nobs];
            ix->minval[k] = ix->minval[nobs];
            ix->maxval[k] = ix->maxval[nobs];
            ix->bits[k] = ix->bits[nobs];
            ++k;
            ix->vals.resize(k);
            ix->minval.resize(k);
            ix->maxval.resize(k);
            ix->bits.resize(k);
        }
    }
    return ix;
}

void write_index(const Index& ix, std::vector<double>& keys, std::vector<int64_t>& offs,
                 std::vector<word_t>& bms) {
    keys.clear(); offs.clear(); bms.clear();
    const uint32_t nobs = (uint32_t)ix.bits.size();
    if (ix.kind == 1) keys = ix.vals;
    else { keys = ix.minval; keys.insert(keys.end(), ix.maxval.begin(), ix.maxval.end()); }
    offs.resize(nobs + 1);
    offs[0] = 0;
    for (uint32_t j = 0; j < nobs; ++j) {
        if (ix.bits[j]) {
            std::vector<word_t> s = ix.bits[j]->serial();
            bms.insert(bms.end(), s.begin(), s.end());
        }
        offs[j + 1] = (int64_t)bms.size();
    }
}

// relic::locate -- src/irelic.cpp:1362-1411, 1416-1809 (LT/LE cases)
static uint32_t relic_locate(const std::vector<double>& vals, double val) {
    const uint32_t nval = (uint32_t)vals.size();
    if (nval == 0) return 0;
    if (!(val >= vals[0])) return 0;
    if (!(val < vals[nval - 1])) return vals[nval - 1] < DBL_MAX ? nval : nval - 1;
    if (nval >= 8) {
        uint32_t i0 = 0, i1 = nval, it = nval / 2;
        while (i0 < it) {
            if (val < vals[it]) i1 = it; else i0 = it;
            it = (i0 + i1) / 2;
        }
        return i1;
    }
    for (uint32_t i = 0; i < nval; ++i)
        if (val < vals[i]) return i;
    return nval;
}

// Bytes of a bitmap as index::sumBins counts them (bitvector::bytes()).
static double bm_bytes(const BV* b) { return b ? 4.0 * b->v.size() + 40.0 : 0.0; }

// index::sumBins(ib, ie, res) -- src/index.cpp:6450-7019, by its result:
// OR of bits[ib, ie), or the complement of the OR of the rest.
static BV sum_bins(const Index& ix, uint32_t ib, uint32_t ie) {
    const uint32_t nobs = (uint32_t)ix.bits.size();
    BV res;
    if (ie > nobs) ie = nobs;
    if (ib >= ie) { res.set(0, ix.nrows); return res; }
    double all = 0, mid = 0;
    for (uint32_t i = 0; i < nobs; ++i) {
        all += bm_bytes(ix.bits[i]);
        if (i >= ib && i < ie) mid += bm_bytes(ix.bits[i]);
    }
    if (mid == 0) { res.set(0, ix.nrows); return res; }
    if (all == mid) { res.set(1, ix.nrows); return res; }
    const bool straight = mid <= std::floor(all / 2);
    res.set(0, ix.nrows);
    auto orin = [&](uint32_t i) { if (ix.bits[i]) { res.op_inplace(*ix.bits[i], OP_OR); res.compress(); } };
    if (straight) {
        for (uint32_t i = ib; i < ie; ++i) orin(i);
    } else {
        for (uint32_t i = 0; i < ib; ++i) orin(i);
        for (uint32_t i = ie; i < nobs; ++i) orin(i);
        res.size();
        res.flip();
    }
    return res;
}

template <class T>
static T at(const void* col, uint32_t i) { return ((const T*)col)[i]; }

static double value_at(const void* col, int type, uint32_t i) {
    switch (type) {
    case COL_I32: return at<int32_t>(col, i);
    case COL_U32: return at<uint32_t>(col, i);
    case COL_U16: return at<uint16_t>(col, i);
    case COL_F32: return at<float>(col, i);
    default: return at<double>(col, i);
    }
}

// part::doScan's integer-bound adjustment (src/part.cpp:11130-11237)
template <class T>
static void int_adjust(double lo, int& lop, double hi, int& rop, T& lb, T& rb) {
    const double tmin = (double)std::numeric_limits<T>::min();
    const double tmax = (double)std::numeric_limits<T>::max();
    lb = 0; rb = 0;
    if (lop == OPLT) {
        if (lo < tmin) { lb = std::numeric_limits<T>::min(); lop = OPLE; }
        else if (lo > tmax) lb = std::numeric_limits<T>::max();
        else lb = (T)lo;
    } else if (lop == OPLE) {
        if (lo < tmin) { lb = std::numeric_limits<T>::min(); lop = OPLE; }
        else if (lo > tmax) { lb = std::numeric_limits<T>::max(); lop = OPLT; }
        else { lb = (T)lo; if ((double)lb != lo) lop = OPLT; }
    }
    if (rop == OPLE) {
        if (hi < tmin) { rb = std::numeric_limits<T>::min(); rop = OPLT; }
        else if (hi > tmax) rb = std::numeric_limits<T>::max();
        else rb = (T)hi;
    } else if (rop == OPLT) {
        if (hi < tmin) rb = std::numeric_limits<T>::min();
        else if (hi > tmax) { rb = std::numeric_limits<T>::max(); rop = OPLE; }
        else { rb = (T)hi; if ((double)rb < hi) rop = OPLE; }
    }
}

template <class T>
static bool int_in(T x, int lop, T lb, int rop, T rb) {
    bool ok = true;
    if (lop == OPLT) ok = x > lb; else if (lop == OPLE) ok = x >= lb;
    if (rop == OPLT) ok = ok && x < rb; else if (rop == OPLE) ok = ok && x <= rb;
    return ok;
}

// the range test of part::doScan for the row
static bool in_range(const void* col, int type, uint32_t i, double lo, int lop, double hi, int rop) {
    int l = lop, r = rop;
    switch (type) {
    case COL_I32: { int32_t a, b; int_adjust<int32_t>(lo, l, hi, r, a, b); return int_in<int32_t>(at<int32_t>(col, i), l, a, r, b); }
    case COL_U32: { uint32_t a, b; int_adjust<uint32_t>(lo, l, hi, r, a, b); return int_in<uint32_t>(at<uint32_t>(col, i), l, a, r, b); }
    case COL_U16: { uint16_t a, b; int_adjust<uint16_t>(lo, l, hi, r, a, b); return int_in<uint16_t>(at<uint16_t>(col, i), l, a, r, b); }
    default: {
        const double x = value_at(col, type, i);   // float widened to double
        bool ok = true;
        if (lop == OPLT) ok = lo < x; else if (lop == OPLE) ok = lo <= x;
        if (rop == OPLT) ok = ok && x < hi; else if (rop == OPLE) ok = ok && x <= hi;
        return ok;
    }
    }
}

BV evaluate(const Index& ix, const void* col, int type, double lo, int lop, double hi, int rop) {
    const uint32_t nobs = (uint32_t)ix.bits.size();
    if (ix.kind == 1) {
        uint32_t bin0 = lop != OPN ? relic_locate(ix.vals, lo) : 0;
        uint32_t bin1 = rop != OPN ? relic_locate(ix.vals, hi) : 0;
        uint32_t hit0, hit1;
        if (lop == OPLT) hit0 = bin0;
        else if (lop == OPLE) hit0 = bin0 > 0 ? bin0 - (lo == ix.vals[bin0 - 1]) : 0;
        else hit0 = 0;
        if (rop == OPLT) hit1 = bin1 > 0 ? bin1 - (hi == ix.vals[bin1 - 1]) : 0;
        else if (rop == OPLE) hit1 = bin1;
        else hit1 = nobs;
        if (hit1 < hit0) hit1 = hit0;
        return sum_bins(ix, hit0, hit1);
    }
    // bin::locate(expr, cand0, cand1, hit0, hit1) -- src/ibin.cpp:7596-8460
    const uint32_t bin0 = lop != OPN ? bin_locate(ix.vals, lo) : 0;
    const uint32_t bin1 = rop != OPN ? bin_locate(ix.vals, hi) : 0;
    uint32_t cand0, cand1, hit0, hit1;
    if (lop == OPLE) {
        if (bin0 >= nobs) { hit0 = cand0 = nobs; }
        else if (lo > ix.maxval[bin0]) { hit0 = cand0 = bin0 + 1; }
        else if (lo <= ix.minval[bin0]) { hit0 = cand0 = bin0; }
        else { hit0 = bin0 + 1; cand0 = bin0; }
    } else if (lop == OPLT) {
        if (bin0 >= nobs) { hit0 = cand0 = nobs; }
        else if (lo >= ix.maxval[bin0]) { hit0 = cand0 = bin0 + 1; }
        else if (lo < ix.minval[bin0]) { hit0 = cand0 = bin0; }
        else { hit0 = bin0 + 1; cand0 = bin0; }
    } else { hit0 = cand0 = 0; }
    if (rop == OPLT) {
        if (bin1 >= nobs) { hit1 = cand1 = nobs; }
        else if (hi > ix.maxval[bin1]) { hit1 = cand1 = bin1 + 1; }
        else if (hi <= ix.minval[bin1]) { hit1 = cand1 = bin1; }
        else { hit1 = bin1; cand1 = bin1 + 1; }
    } else if (rop == OPLE) {
        if (bin1 >= nobs) { hit1 = cand1 = nobs; }
        else if (hi >= ix.maxval[bin1]) { hit1 = cand1 = bin1 + 1; }
        else if (hi < ix.minval[bin1]) { hit1 = cand1 = bin1; }
        else { hit1 = bin1; cand1 = bin1 + 1; }
    } else { hit1 = cand1 = nobs; }
    if (hit1 < hit0) hit1 = hit0;
    BV lower = sum_bins(ix, hit0, hit1);
    // edge bins through the raw values (checkBin -> doScan fallback,
    // src/ibin.cpp:6651-6704)
    auto check = [&](uint32_t b) {
        if (!ix.bits[b]) return;
        std::vector<word_t> r;
        ix.bits[b]->rids(r);
        BV delta;
        for (word_t row : r)
            if (in_range(col, type, row, lo, lop, hi, rop)) delta.setBit(row, 1);
        delta.adjustSize(0, ix.nrows);
        lower.op_inplace(delta, OP_OR);
        lower.compress();
    };
    if (cand0 < hit0) check(cand0);
    if (cand1 > hit1 && hit1 < nobs) check(hit1);
    return lower;
}

} // namespace orc
