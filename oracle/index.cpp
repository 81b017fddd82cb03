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
    ix->offsets_built = true;     // offset64 filled at src/irelic.cpp:186-190
    return ix;
}

// index::optionalUnpack without a <compressing> directive
// (src/index.cpp:8803-8854): bitmaps whose bytes() (code words * 4 +
// sizeof(bitvector), 40) exceed max(nrows/9, 52) are decompressed.
static void optional_unpack(Index* ix) {
    const uint64_t bar1 = std::max<uint64_t>(ix->nrows / 9u, 40u + 12u);
    for (BV* b : ix->bits)
        if (b != nullptr && 4ull * b->v.size() + 40u > bar1) b->decompress();
}

// src/index.cpp:2552-2611 (a std::map from value to bitvector, setBit per
// row), then every bitmap padded to nrows; relic::construct ends with
// optionalUnpack (src/irelic.cpp:910)
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
    optional_unpack(ix);
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

// src/ibin.cpp:3350-3449, then optionalUnpack
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
    optional_unpack(ix);   // bin::construct, src/ibin.cpp:2960
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

// bitvector::getSerialSize() in bytes (src/bitvector.h:150-154)
static double serial_bytes(const Index& ix, uint32_t b) {
    if (!ix.bits[b]) return 0.0;
    return 4.0 * (double)(ix.bits[b]->v.size() + 1 + (ix.nrows % 31u ? 1u : 0u));
}

// index::sumBins(ib, ie, res) -- src/index.cpp:6450-7019, by its result:
// OR of bits[ib, ie), or the complement of the OR of the rest.
static BV sum_bins(const Index& ix, uint32_t ib, uint32_t ie) {
    const uint32_t nobs = (uint32_t)ix.bits.size();
    BV res;
    if (ie > nobs) ie = nobs;
    if (ib >= ie) { res.set(0, ix.nrows); return res; }
    // the measure is offset64 (src/index.cpp:6463-6477): the serialized
    // sizes a constructor or a file put there, else bitvector::bytes()
    double all = 0, mid = 0;
    for (uint32_t i = 0; i < nobs; ++i) {
        const double m = ix.offsets_built ? serial_bytes(ix, i) : bm_bytes(ix.bits[i]);
        all += m;
        if (i >= ib && i < ie) mid += m;
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

// relic::locate(expr, hit0, hit1) -- src/irelic.cpp:1416-1809 (the LT/LE
// cases of the two sides)
static void relic_hits(const Index& ix, double lo, int lop, double hi, int rop, uint32_t& hit0,
                       uint32_t& hit1) {
    const uint32_t nobs = (uint32_t)ix.bits.size();
    const uint32_t bin0 = lop != OPN ? relic_locate(ix.vals, lo) : 0;
    const uint32_t bin1 = rop != OPN ? relic_locate(ix.vals, hi) : 0;
    if (lop == OPLT) hit0 = bin0;
    else if (lop == OPLE) hit0 = bin0 > 0 ? bin0 - (lo == ix.vals[bin0 - 1]) : 0;
    else hit0 = 0;
    if (rop == OPLT) hit1 = bin1 > 0 ? bin1 - (hi == ix.vals[bin1 - 1]) : 0;
    else if (rop == OPLE) hit1 = bin1;
    else hit1 = nobs;
    if (hit1 < hit0) hit1 = hit0;
}

// bin::locate(expr, cand0, cand1, hit0, hit1) -- src/ibin.cpp:7596-8460
static void bin_range(const Index& ix, double lo, int lop, double hi, int rop, uint32_t& cand0,
                      uint32_t& cand1, uint32_t& hit0, uint32_t& hit1) {
    const uint32_t nobs = (uint32_t)ix.bits.size();
    const uint32_t bin0 = lop != OPN ? bin_locate(ix.vals, lo) : 0;
    const uint32_t bin1 = rop != OPN ? bin_locate(ix.vals, hi) : 0;
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
}

BV evaluate(const Index& ix, const void* col, int type, double lo, int lop, double hi, int rop) {
    const uint32_t nobs = (uint32_t)ix.bits.size();
    if (ix.kind == 1) {
        uint32_t hit0, hit1;
        relic_hits(ix, lo, lop, hi, rop, hit0, hit1);
        return sum_bins(ix, hit0, hit1);
    }
    uint32_t cand0, cand1, hit0, hit1;
    bin_range(ix, lo, lop, hi, rop, cand0, cand1, hit0, hit1);
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

static uint32_t elem_size(int type);

// src/ibin.cpp:6722-6829.  The cost test reads offset64: the serialized
// sizes where a file or the keys/offsets constructor filled it, else the
// bitvector::bytes() that sumBins fills it with (src/index.cpp:6463-6477);
// the state of an index that has evaluated a range (the reference's first
// estimate on a fresh in-memory index, before any sumBins, sees no
// offsets and keeps only the edge term).
void estimate(const Index& ix, const void* col, int type, double lo, int lop, double hi, int rop,
              BV& lower, BV& upper) {
    upper.clear();
    if (ix.kind == 1) {
        lower = evaluate(ix, col, type, lo, lop, hi, rop);
        return;
    }
    const uint32_t nobs = (uint32_t)ix.bits.size();
    if (nobs == 0 || ix.nrows == 0) { lower.set(0, ix.nrows); return; }
    uint32_t cand0, cand1, hit0, hit1;
    bin_range(ix, lo, lop, hi, rop, cand0, cand1, hit0, hit1);
    double cost = 0.0;
    if (cand0 < cand1 && cand1 < nobs + 1) {
        std::vector<int64_t> off(nobs + 1, 0);
        for (uint32_t i = 0; i < nobs; ++i)
            off[i + 1] = off[i] + (int64_t)(ix.offsets_built ? serial_bytes(ix, i) : bm_bytes(ix.bits[i]));
        const int64_t tot = off[nobs] - off[0];
        const int64_t mid = off[cand1] - off[cand0];
        cost = (tot >> 1) >= mid ? (double)mid : (double)(tot - mid);
    }
    const double edge = (double)elem_size(type) * (double)ix.nrows / nobs;
    if (hit0 > cand0 && cand1 > hit1) cost += edge * 2.0;
    else if (hit0 > cand0 || cand1 > hit1) cost += edge;
    if (cand0 >= cand1) {
        lower.set(0, ix.nrows);
        upper.set(0, ix.nrows);
    } else if (cost > ix.nrows * 0.75) {
        lower.set(0, ix.nrows);
        upper.set(1, ix.nrows);
    } else if (hit0 < hit1) {
        lower = sum_bins(ix, hit0, hit1);
        if (cand0 < hit0 || (cand1 > hit1 && hit1 < nobs)) {
            upper = lower;
            if (cand0 < hit0 && ix.bits[cand0]) { upper.op_inplace(*ix.bits[cand0], OP_OR); upper.compress(); }
            if (cand1 > hit1 && hit1 < nobs && ix.bits[hit1]) { upper.op_inplace(*ix.bits[hit1], OP_OR); upper.compress(); }
        }
    } else {
        lower.set(0, ix.nrows);
        upper = sum_bins(ix, cand0, cand1);
    }
}

// part::doScan over every row (src/part.cpp:11130-11237 for integer
// columns, 12094-12219 / 12849- for float / double).
BV scan(const void* col, int type, uint32_t n, double lo, int lop, double hi, int rop) {
    BV r;
    for (uint32_t i = 0; i < n; ++i)
        if (in_range(col, type, i, lo, lop, hi, rop)) r.setBit(i, 1);
    r.adjustSize(0, n);
    return r;
}

static uint32_t elem_size(int type) {
    return type == COL_U16 ? 2u : type == COL_F64 ? 8u : 4u;
}

double estimate_cost(const Index& ix, int type, double lo, int lop, double hi, int rop, int model) {
    const uint32_t nobs = (uint32_t)ix.bits.size();
    // offset64 is filled by relic(col, card, ind) and by reading an index
    // file; the in-memory constructors leave it empty
    const bool offsets = model == COST_PARTITION || ix.offsets_built;
    auto span = [&](uint32_t b0, uint32_t b1) {
        double m = 0;
        for (uint32_t b = b0; b < b1 && b < nobs; ++b) m += serial_bytes(ix, b);
        return m;
    };
    if (ix.kind == 1) {
        // relic::estimateCost -- src/irelic.cpp:1857-1902
        uint32_t h0, h1;
        relic_hits(ix, lo, lop, hi, rop, h0, h1);
        if (h0 >= h1) return 0.0;
        if (offsets) {
            if (h1 > h0 + 1) {
                const double tot = span(0, nobs), mid = span(h0, h1);
                return std::floor(tot / 2) >= mid ? mid : tot - mid;
            }
            return 0.5 * span(h0, h1);
        }
        if (h1 > h0 + 1) {
            double r = 0;
            for (uint32_t b = h0; b < h1 && b < nobs; ++b) r += bm_bytes(ix.bits[b]);
            return r;
        }
        return h0 < nobs ? 0.5 * bm_bytes(ix.bits[h0]) : 0.0;
    }
    // bin::estimateCost -- src/ibin.cpp:6531-6569
    double r = 0;
    uint32_t cand0 = 0, cand1 = nobs, hit0 = nobs, hit1 = 0;
    if (offsets) {
        bin_range(ix, lo, lop, hi, rop, cand0, cand1, hit0, hit1);
        if (cand0 < cand1 && cand1 <= nobs) {
            const double tot = span(0, nobs), mid = span(cand0, cand1);
            r = std::floor(tot / 2) >= mid ? mid : tot - mid;
        }
    }
    if (hit0 > cand0 || hit1 < cand1) {
        const uint32_t es = elem_size(type);
        if (nobs > 0) {
            // elementSize() * nrows / nobs in unsigned 32-bit arithmetic
            const double ccheck = (double)((uint32_t)(es * ix.nrows) / nobs);
            r += (hit0 > cand0 && hit1 < cand1 && hit0 <= hit1) ? 2.0 * ccheck : ccheck;
        } else {
            r += (double)(uint32_t)(es * ix.nrows);
        }
    }
    return r;
}

// part::countPages over a mask of n ones (src/part.cpp:2616-2678): the
// indexSet walk sees [0, 31g) as one range, then the active word's rows.
static uint32_t full_mask_pages(uint32_t n, uint32_t es) {
    if (n == 0) return 0;
    const uint32_t wpp = 4096u / es;
    const uint32_t g = n / 31u, anb = n % 31u;
    uint32_t res = 0, last = 0;
    if (g > 0) { res += (31u * g) / wpp; last = 31u * g; }
    if (anb > 0) {
        const uint32_t first = 31u * g;
        const uint32_t p0 = first / wpp;
        res += (last < p0 * wpp) ? 1u : 0u;
        last = n - 1;
        res += (last / wpp > p0) ? 1u : 0u;
    }
    return res == 0 ? 1u : res;
}

double scan_cost(const Index& ix, int type, int model) {
    if (model == COST_PARTITION) {
        // column::evaluateRange -- src/column.cpp:6102-6111, pageSize 4096
        return 4096.0 * full_mask_pages(ix.nrows, elem_size(type)) + 8.0 * ix.nrows / 4096.0;
    }
    // bord::column::evaluateRange -- src/bord.cpp:5903-5912
    return ix.nrows * 0.5 + 999.0;
}

// qContinuousRange::overlap (src/qExpr.cpp:1673-1803) for lo OP x OP hi
// against the column's [cmin, cmax]
static bool overlap(double lo, int lop, double hi, int rop, double cmin, double cmax) {
    if (!(cmin <= cmax)) return true;
    if (lop == OPLT && !(lo < cmax)) return false;
    if (lop == OPLE && !(lo <= cmax)) return false;
    if (lop == OPN) {
        if (rop == OPLT) return cmin < hi;
        if (rop == OPLE) return cmin <= hi;
        return false;
    }
    if (rop == OPLT) return cmin < hi;
    if (rop == OPLE) return cmin <= hi;
    return true;
}

int auto_plan(const Index& ix, const void* col, int type, double lo, int lop, double hi, int rop,
              int model, bool* empty) {
    *empty = false;
    if (model == COST_PARTITION && ix.nrows > 0) {
        double cmin = DBL_MAX, cmax = -DBL_MAX;
        for (uint32_t i = 0; i < ix.nrows; ++i) {
            const double x = value_at(col, type, i);
            if (x < cmin) cmin = x;
            if (x > cmax) cmax = x;
        }
        if (!overlap(lo, lop, hi, rop, cmin, cmax)) { *empty = true; return PLAN_INDEX; }
    }
    const double ic = estimate_cost(ix, type, lo, lop, hi, rop, model);
    return ic < scan_cost(ix, type, model) ? PLAN_INDEX : PLAN_SCAN;
}

BV evaluate_plan(const Index& ix, const void* col, int type, double lo, int lop, double hi, int rop,
                 int plan, int model, int* used) {
    bool empty = false;
    if (plan == 0) plan = auto_plan(ix, col, type, lo, lop, hi, rop, model, &empty);
    if (used) *used = plan;
    if (lop == OPN && rop == OPN) {          // no bound: every row
        BV r;
        r.set(1, ix.nrows);
        return r;
    }
    if (empty) {
        BV r;
        r.set(0, ix.nrows);
        return r;
    }
    if (plan == PLAN_SCAN) return scan(col, type, ix.nrows, lo, lop, hi, rop);
    return evaluate(ix, col, type, lo, lop, hi, rop);
}

// ---------------------------------------------------------------------------
// Two-level encodings (ibis::fuzz, ibis::zona).

// array_t<T>::find (src/array_t.cpp:507-528): the first position whose
// value is not less than v.
template <class T>
static uint32_t find_first(const std::vector<T>& a, T v) {
    return (uint32_t)(std::lower_bound(a.begin(), a.end(), v) - a.begin());
}

// The boundaries of the coarse bins: fine bitmaps split into groups of
// nearly equal offset64 bytes (src/ixfuzz.cpp:215-231, src/ixzona.cpp:
// 146-161; fuzz keeps the target in 32 bits).
static std::vector<uint32_t> coarse_bounds(const std::vector<int64_t>& off, uint32_t ncoarse, bool i32) {
    const uint32_t nbits = (uint32_t)off.size() - 1;
    std::vector<uint32_t> cb(ncoarse + 1);
    cb[0] = 0;
    for (uint32_t i = 1; i < ncoarse; ++i) {
        int64_t target = off[cb[i - 1]] + (off.back() - off[cb[i - 1]]) / (int64_t)(ncoarse - i + 1);
        if (i32) target = (int32_t)target;
        cb[i] = find_first(off, target);
        if (cb[i] > cb[i - 1] + 1 && off[cb[i]] - target > target - off[cb[i] - 1]) --cb[i];
        else if (cb[i] <= cb[i - 1]) cb[i] = cb[i - 1] + 1;
    }
    cb[ncoarse] = nbits;
    for (uint32_t i = ncoarse - 1; i > 0 && cb[i + 1] < cb[i]; --i) cb[i] = cb[i + 1] - 1;
    return cb;
}

// OR of fine bitmaps [a, b) as a compressed bitvector of nrows bits.
static BV or_range(const Index& ix, uint32_t a, uint32_t b) {
    BV r;
    r.set(0, ix.nrows);
    for (uint32_t i = a; i < b && i < ix.bits.size(); ++i)
        if (ix.bits[i]) { r.op_inplace(*ix.bits[i], OP_OR); r.compress(); }
    return r;
}

TwoLevel* coarsen(const Index& ix, int enc, uint32_t ncoarse) {
    TwoLevel* tl = new TwoLevel();
    tl->enc = enc;
    const uint32_t nbits = (uint32_t)ix.bits.size();
    // offset64 (src/ixfuzz.cpp:166-180): the file's positions where the
    // index has them, else bitvector::bytes() of every bitmap
    tl->off.assign(nbits + 1, 0);
    for (uint32_t i = 0; i < nbits; ++i)
        tl->off[i + 1] = tl->off[i] + (int64_t)(ix.offsets_built ? serial_bytes(ix, i) : bm_bytes(ix.bits[i]));
    if (ix.vals.size() < 32) return tl;              // no coarse level for fewer than 32 values
    const uint32_t nrows = ix.nrows;
    if (enc == ENC_FUZZ) {
        // src/ixfuzz.cpp:182-208: sf = fine-level words; the default count
        // minimises the product of index size and query cost
        if (ncoarse < 5u && tl->off.back() > tl->off[0] + (int64_t)(nrows / 31u)) {
            const int wm1 = 31;
            const long sf = (long)((tl->off.back() - tl->off[0]) / 4);
            unsigned nc = (unsigned)(wm1 * sf / (std::sqrt(2.0) * nrows));
            const unsigned ncmax = (unsigned)std::sqrt(2.0 * ix.vals.size());
            if (nc < ncmax) {
                const double obj1 = (sf + (nc + 1 - std::ceil(0.5 * nc)) * nrows / wm1) *
                                    (sf * 0.5 / nc + 2.0 * nrows / wm1);
                const double obj2 = (sf + (nc + 2 - std::ceil(0.5 * nc + 0.5)) * nrows / wm1) *
                                    (sf * 0.5 / (nc + 1.0) + 2.0 * nrows / wm1);
                nc += (obj2 < obj1);
            } else {
                nc = ncmax;
            }
            ncoarse = nc;
        }
        if (ncoarse < 5 || ncoarse >= ix.vals.size()) return tl;
    } else {
        ncoarse = 11;                                 // src/ixzona.cpp:128
    }
    tl->cbounds = coarse_bounds(tl->off, ncoarse, enc == ENC_FUZZ);
    const std::vector<uint32_t>& cb = tl->cbounds;
    if (enc == ENC_FUZZ) {
        // interval encoding (src/ixfuzz.cpp:242-253): bitmap i covers the
        // nc2 coarse bins from i
        const uint32_t nc2 = (ncoarse + 1) / 2, ncb = ncoarse - nc2 + 1;
        for (uint32_t i = 0; i < ncb; ++i) tl->cbits.push_back(new BV(or_range(ix, cb[i], cb[i + nc2])));
    } else {
        for (uint32_t i = 0; i < ncoarse; ++i) tl->cbits.push_back(new BV(or_range(ix, cb[i], cb[i + 1])));
    }
    tl->coff.assign(tl->cbits.size() + 1, 0);
    for (size_t i = 0; i < tl->cbits.size(); ++i) {
        tl->cbits[i]->compress();
        tl->coff[i + 1] = tl->coff[i] + (int64_t)bm_bytes(tl->cbits[i]);
    }
    return tl;
}

TwoLevel* two_level_from_file(const Index& ix, int enc, const std::vector<int64_t>& fine_pos,
                              const std::vector<uint32_t>& cbounds, const std::vector<int64_t>& coff,
                              const std::vector<std::vector<word_t>>& cbits_serial) {
    (void)ix;
    TwoLevel* tl = new TwoLevel();
    tl->enc = enc;
    tl->off = fine_pos;
    tl->cbounds = cbounds;
    tl->coff = coff;
    for (const auto& w : cbits_serial) tl->cbits.push_back(new BV(BV::from_serial(w.data(), w.size())));
    return tl;
}

namespace {

// One evaluation: the fine level, the coarse level, and the bitmaps read.
struct TLEval {
    const Index& ix;
    const TwoLevel& tl;
    uint64_t loads = 0;
    const std::vector<int64_t>& off;
    const std::vector<int64_t>& coff;
    const std::vector<uint32_t>& cb;
    TLEval(const Index& i, const TwoLevel& t) : ix(i), tl(t), off(t.off), coff(t.coff), cb(t.cbounds) {}
    uint32_t nobs() const { return (uint32_t)ix.bits.size(); }
    BV none() const { BV r; r.set(0, ix.nrows); return r; }
    BV all() const { BV r; r.set(1, ix.nrows); return r; }
    const BV* coarse(uint32_t i) {
        if (i >= tl.cbits.size()) return nullptr;
        loads |= 1ull << i;
        return tl.cbits[i];
    }
    void load_range(uint32_t i, uint32_t j) {
        for (; i < j && i < tl.cbits.size(); ++i) loads |= 1ull << i;
    }
    // the fine level's cost of [h0, h1): the smaller of the range and its
    // complement (src/ixfuzz.cpp:793-804)
    long fine_cost(uint32_t h0, uint32_t h1) const {
        const long in = (long)(off[h1] - off[h0]);
        const long out = (long)((off.back() - off[h1]) + (off[h0] - off[0]));
        return in <= out ? in : out;
    }
    // index::addBins (src/index.cpp:5537-5580): OR of [ib, ie) into res;
    // the whole fine level sets every row
    void add_bins(uint32_t ib, uint32_t ie, BV& res) const {
        if (ie > nobs()) ie = nobs();
        if (ib >= ie) return;
        if (ib == 0 && ie == nobs()) { res = all(); return; }
        BV r = or_range(ix, ib, ie);
        res.op_inplace(r, OP_OR);
        res.compress();
    }
    void minus(BV& res, const BV& b) const { res.op_inplace(b, OP_MINUS); res.compress(); }
    void andw(BV& res, const BV& b) const { res.op_inplace(b, OP_AND); res.compress(); }
    void orw(BV& res, const BV& b) const { res.op_inplace(b, OP_OR); res.compress(); }

    // fuzz::coarseEstimate (src/ixfuzz.cpp:626-674): bytes of the coarse
    // bitmaps coarseEvaluate reads for coarse bins [lo, hi)
    long fuzz_coarse_cost(uint32_t lo, uint32_t hi) const {
        const uint32_t mid = (uint32_t)cb.size() / 2;
        auto sz = [&](uint32_t i) { return (long)(coff[i + 1] - coff[i]); };
        if (lo >= cb.size() || lo >= hi) return 0;
        if (hi > mid) {
            long c = sz(hi - mid);
            if (lo > hi - mid) c += lo >= mid ? sz(lo - mid) : sz(lo);
            else if (lo < hi - mid) c += sz(lo);
            return c;
        }
        if (hi < mid) return sz(lo) + sz(hi);
        return sz(0) + (lo > 0 ? sz(lo) : 0);
    }
    // fuzz::coarseEvaluate (src/ixfuzz.cpp:676-765): rows of coarse bins
    // [lo, hi) from at most two interval bitmaps
    BV fuzz_coarse(uint32_t lo, uint32_t hi) {
        const uint32_t mid = (uint32_t)cb.size() / 2;
        const uint32_t ncb = (uint32_t)tl.cbits.size();
        if (lo >= cb.size() || lo >= hi) return none();
        if (lo + 1 == hi) {
            const uint32_t a = hi < ncb ? lo : hi - mid, b = hi < ncb ? hi : lo - mid;
            load_range(std::min(a, b), std::min(a, b) + 2);
            const BV* x = tl.cbits[a];
            if (!x) return none();
            BV r = *x;
            if (tl.cbits[b]) minus(r, *tl.cbits[b]);
            return r;
        }
        if (hi > mid) {
            const BV* x = coarse(hi - mid);
            BV r = x ? *x : none();
            if (lo > hi - mid) {
                if (lo >= mid) {
                    const BV* y = coarse(lo - mid);
                    if (y) minus(r, *y);
                } else {
                    const BV* y = coarse(lo);
                    if (y) andw(r, *y); else r = none();
                }
            } else if (lo < hi - mid) {
                const BV* y = coarse(lo);
                if (y) orw(r, *y);
            }
            return r;
        }
        if (hi < mid) {
            const BV* x = coarse(lo);
            if (!x) return none();
            BV r = *x;
            const BV* y = coarse(hi);
            if (y) minus(r, *y);
            return r;
        }
        const BV* x = coarse(0);
        BV r = x ? *x : none();
        if (lo > 0) {
            const BV* y = coarse(lo);
            if (y) andw(r, *y);
        }
        return r;
    }
    // zona: OR of coarse bitmaps [a, b) (index::addBits) read straight, or
    // the rows outside [a, b) complemented, whichever is smaller
    // (src/ixzona.cpp:889-913)
    BV zona_coarse(uint32_t a, uint32_t b, uint32_t out_lo_end, uint32_t out_hi_begin, uint32_t ncoarse,
                   bool lo_cond, bool hi_cond) {
        const long csize = (long)(coff.back() - coff[0]);
        const long tmp = (long)(coff[b] - coff[a]);
        BV r;
        if (tmp + tmp <= csize) {
            r = none();
            load_range(a, b);
            for (uint32_t i = a; i < b; ++i)
                if (tl.cbits[i]) orw(r, *tl.cbits[i]);
        } else {
            BV bv = none();
            if (lo_cond) {
                load_range(0, out_lo_end);
                for (uint32_t i = 0; i < out_lo_end; ++i)
                    if (tl.cbits[i]) orw(bv, *tl.cbits[i]);
            }
            if (hi_cond) {
                load_range(out_hi_begin, ncoarse);
                for (uint32_t i = out_hi_begin; i < ncoarse; ++i)
                    if (tl.cbits[i]) orw(bv, *tl.cbits[i]);
            }
            r = all();
            minus(r, bv);
        }
        return r;
    }
    long zona_part(uint32_t a, uint32_t b) const {
        const long csize = (long)(coff.back() - coff[0]);
        const long t = (long)(coff[b] - coff[a]);
        return t + t <= csize ? t : csize - t;
    }
};

} // namespace

double two_level_cost(const Index& ix, const TwoLevel& tl, int type, double lo, int lop, double hi, int rop) {
    TLEval E(ix, tl);
    const uint32_t nobs = E.nobs();
    const double scan = (double)elem_size(type) * ix.nrows;
    if (nobs == 0 || tl.off.size() != nobs + 1) return scan;
    uint32_t h0 = 0, h1 = 0;
    relic_hits(ix, lo, lop, hi, rop, h0, h1);
    const auto& off = tl.off;
    const auto& coff = tl.coff;
    const auto& cb = tl.cbounds;
    const bool have_coff = !tl.cbits.empty() && coff.size() == tl.cbits.size() + 1;
    if (tl.enc == ENC_FUZZ) {
        if (h1 <= h0 || h0 >= nobs) return 0.0;
        if (h0 == 0 && h1 >= nobs) return 0.0;
        const uint32_t ncoarse = cb.empty() ? 0u : (uint32_t)cb.size() - 1;
        const long fine = E.fine_cost(h0, h1);
        if (h0 + 3 >= h1 || ncoarse == 0 || !have_coff) return (double)fine;
        const uint32_t c0 = find_first(cb, h0), c1 = find_first(cb, h1);
        if (c0 >= c1) {
            const long tmp = E.fuzz_coarse_cost(c1 - 1, c1) + (long)((off[h0] - off[cb[c1 - 1]]) + (off[cb[c1]] - off[h1]));
            return (double)(tmp / 100 >= fine / 99 ? fine : tmp);
        }
        long cost = E.fuzz_coarse_cost(c0, c1 - 1) + (long)((off[cb[c0]] - off[h0]) + (off[h1] - off[cb[c1 - 1]]));
        if (c0 > 0)
            cost = std::min(cost, E.fuzz_coarse_cost(c0 - 1, c1 - 1) +
                                      (long)((off[h0] - off[cb[c0 - 1]]) + (off[h1] - off[cb[c1 - 1]])));
        cost = std::min(cost, E.fuzz_coarse_cost(c0, c1) + (long)((off[cb[c0]] - off[h0]) + (off[cb[c1]] - off[h1])));
        if (c0 > 0)
            cost = std::min(cost, E.fuzz_coarse_cost(c0 - 1, c1) +
                                      (long)((off[h0] - off[cb[c0 - 1]]) + (off[cb[c1]] - off[h1])));
        return (double)(cost / 100 >= fine / 99 ? fine : cost);
    }
    // zona
    if (h1 <= h0) return 0.0;
    if (h0 == 0 && h1 >= nobs) return 0.0;
    const long fine = E.fine_cost(h0, h1);
    const uint32_t ncoarse = (tl.cbits.empty() || cb.empty()) ? 0u
                             : tl.cbits.size() + 1 <= cb.size() ? (uint32_t)tl.cbits.size() : (uint32_t)cb.size() - 1;
    if (h0 + 1 == h1) return (double)fine;
    if (h0 + 3 >= h1 || !have_coff) return (double)fine;
    const uint32_t c0 = find_first(cb, h0), c1 = find_first(cb, h1);
    if (c0 >= c1) {
        const long tmp = (long)(coff[c0] - coff[c0 - 1]) + (long)(off[cb[c0]] - off[cb[c0 - 1]]) - fine;
        return (double)((long)(0.99 * fine) >= tmp ? tmp : fine);
    }
    if (c1 < ncoarse && cb[c1] == h1 && cb[c0] == h0) return (double)E.zona_part(c0, c1);
    long cost = E.zona_part(c0, c1 - 1) + (long)((off[cb[c0]] - off[h0]) + (off[h1] - off[cb[c1 - 1]]));
    if (c0 > 0)
        cost = std::min(cost, E.zona_part(c0 - 1, c1 - 1) + (long)((off[h0] - off[cb[c0 - 1]]) + (off[h1] - off[cb[c1 - 1]])));
    cost = std::min(cost, E.zona_part(c0, c1) + (long)((off[cb[c0]] - off[h0]) + (off[cb[c1]] - off[h1])));
    if (c0 > 0)
        cost = std::min(cost, E.zona_part(c0 - 1, c1) + (long)((off[h0] - off[cb[c0 - 1]]) + (off[cb[c1]] - off[h1])));
    if (cost > (long)(0.99 * fine)) cost = fine;
    return (double)cost;
}

BV two_level_eval(const Index& ix, const TwoLevel& tl, double lo, int lop, double hi, int rop, int* option,
                  uint64_t* loads) {
    TLEval E(ix, tl);
    const uint32_t nobs = E.nobs();
    *loads = 0;
    *option = TL_EMPTY;
    if (nobs == 0) return E.none();
    uint32_t h0 = 0, h1 = 0;
    relic_hits(ix, lo, lop, hi, rop, h0, h1);
    if (h1 <= h0 || (tl.enc == ENC_FUZZ && h0 >= nobs)) return E.none();
    if (h0 == 0 && h1 >= nobs) { *option = TL_ALL; return E.all(); }
    if (h0 + 1 == h1) {
        *option = TL_EQUAL;
        return ix.bits[h0] ? *ix.bits[h0] : E.none();
    }
    const auto& off = tl.off;
    const auto& coff = tl.coff;
    const auto& cb = tl.cbounds;
    const bool have_coff = !tl.cbits.empty() && coff.size() == tl.cbits.size() + 1;
    const uint32_t ncoarse = tl.enc == ENC_FUZZ
        ? (cb.empty() ? 0u : (uint32_t)cb.size() - 1)
        : ((tl.cbits.empty() || cb.empty()) ? 0u
           : tl.cbits.size() + 1 <= cb.size() ? (uint32_t)tl.cbits.size() : (uint32_t)cb.size() - 1);
    if (h0 + 3 >= h1 || ncoarse == 0 || !have_coff) {
        *option = TL_FINE;
        return sum_bins(ix, h0, h1);
    }
    const uint32_t c0 = find_first(cb, h0), c1 = find_first(cb, h1);
    BV lower;
    if (tl.enc == ENC_FUZZ) {
        const long finec = E.fine_cost(h0, h1);
        if (c0 >= c1) {
            const long tmp = E.fuzz_coarse_cost(c1 - 1, c1) + (long)((off[h0] - off[cb[c1 - 1]]) + (off[cb[c1]] - off[h1]));
            if (finec <= (long)(0.99 * tmp)) { *option = TL_FINE; lower = sum_bins(ix, h0, h1); }
            else {
                *option = TL_SAME_BIN;
                lower = E.fuzz_coarse(c1 - 1, c1);
                if (h0 > cb[c1 - 1]) E.minus(lower, sum_bins(ix, cb[c1 - 1], h0));
                if (cb[c1] > h1) E.minus(lower, sum_bins(ix, h1, cb[c1]));
            }
            *loads = E.loads;
            return lower;
        }
        // the five options (src/ixfuzz.cpp:943-1011); option 4's coarse
        // cost starts at c0-1 here, where estimateCost starts it at c0
        int opt = 2;
        long cost = E.fuzz_coarse_cost(c0, c1 - 1) + (long)((off[cb[c0]] - off[h0]) + (off[h1] - off[cb[c1 - 1]]));
        long tmp;
        if (c0 > 0) {
            tmp = E.fuzz_coarse_cost(c0 - 1, c1 - 1) + (long)((off[h0] - off[cb[c0 - 1]]) + (off[h1] - off[cb[c1 - 1]]));
            if (tmp < cost) { cost = tmp; opt = 3; }
        }
        tmp = E.fuzz_coarse_cost(c0 > 0 ? c0 - 1 : 0, c1) + (long)((off[cb[c0]] - off[h0]) + (off[cb[c1]] - off[h1]));
        if (tmp < cost) { cost = tmp; opt = 4; }
        if (c0 > 0) {
            tmp = E.fuzz_coarse_cost(c0 - 1, c1) + (long)((off[h0] - off[cb[c0 - 1]]) + (off[cb[c1]] - off[h1]));
            if (tmp < cost) { cost = tmp; opt = 5; }
        }
        if (cost > (long)(0.99 * finec)) opt = 1;
        *option = opt;
        switch (opt) {
        case 2:
            lower = E.fuzz_coarse(c0, c1 - 1);
            if (h0 < cb[c0]) E.add_bins(h0, cb[c0], lower);
            if (cb[c1 - 1] < h1) E.add_bins(cb[c1 - 1], h1, lower);
            break;
        case 3:
            lower = E.fuzz_coarse(c0 - 1, c1 - 1);
            if (cb[c0 - 1] < h0) E.minus(lower, sum_bins(ix, cb[c0 - 1], h0));
            if (cb[c1 - 1] < h1) E.add_bins(cb[c1 - 1], h1, lower);
            break;
        case 4:
            lower = E.fuzz_coarse(c0, c1);
            if (h0 < cb[c0]) E.add_bins(h0, cb[c0], lower);
            if (cb[c1] > h1) E.minus(lower, sum_bins(ix, h1, cb[c1]));
            break;
        case 5:
            lower = E.fuzz_coarse(c0 - 1, c1);
            if (h0 > cb[c0 - 1]) E.minus(lower, sum_bins(ix, cb[c0 - 1], h0));
            if (cb[c1] > h1) E.minus(lower, sum_bins(ix, h1, cb[c1]));
            break;
        default:
            lower = sum_bins(ix, h0, h1);
        }
        *loads = E.loads;
        return lower;
    }
    // zona
    if (c0 >= c1) {
        const long fine = (long)(off[h1] - off[h0]);
        const long tmp = (long)(coff[c0] - coff[c0 - 1]) + (long)(off[cb[c0]] - off[cb[c0 - 1]]) - fine;
        if ((long)(0.99 * fine) <= tmp) { *option = TL_FINE; return sum_bins(ix, h0, h1); }
        *option = TL_SAME_BIN;
        const BV* x = E.coarse(c0 - 1);
        lower = x ? *x : E.all();
        if (h0 > cb[c0 - 1]) E.minus(lower, sum_bins(ix, cb[c0 - 1], h0));
        if (h1 < cb[c0]) E.minus(lower, sum_bins(ix, h1, cb[c0]));
        *loads = E.loads;
        return lower;
    }
    // five pairs (src/ixzona.cpp:803-882)
    int opt = 2;
    long cost = E.zona_part(c0, c1 - 1) + (long)((off[cb[c0]] - off[h0]) + (off[h1] - off[cb[c1 - 1]]));
    long tmp;
    if (c0 > 0) {
        tmp = E.zona_part(c0 - 1, c1 - 1) + (long)((off[h0] - off[cb[c0 - 1]]) + (off[h1] - off[cb[c1 - 1]]));
        if (tmp < cost) { cost = tmp; opt = 3; }
    }
    tmp = E.zona_part(c0, c1) + (long)((off[cb[c0]] - off[h0]) + (off[cb[c1]] - off[h1]));
    if (tmp < cost) { cost = tmp; opt = 4; }
    if (c0 > 0) {
        tmp = E.zona_part(c0 - 1, c1) + (long)((off[h0] - off[cb[c0 - 1]]) + (off[cb[c1]] - off[h1]));
        if (tmp < cost) { cost = tmp; opt = 5; }
    }
    if (cost > (long)(0.99 * E.fine_cost(h0, h1))) opt = 1;
    *option = opt;
    switch (opt) {
    case 2:
        lower = E.none();
        if (c0 < c1 - 1) lower = E.zona_coarse(c0, c1 - 1, c0, c1 - 1, ncoarse, c0 > 0, c1 <= ncoarse);
        if (h0 < cb[c0]) E.add_bins(h0, cb[c0], lower);
        if (cb[c1 - 1] < h1) E.add_bins(cb[c1 - 1], h1, lower);
        break;
    case 3:
        lower = E.zona_coarse(c0 - 1, c1 - 1, c0 - 1, c1 - 1, ncoarse, c0 > 1, c1 <= ncoarse);
        if (cb[c0 - 1] < h0) E.minus(lower, sum_bins(ix, cb[c0 - 1], h0));
        if (cb[c1 - 1] < h1) E.add_bins(cb[c1 - 1], h1, lower);
        break;
    case 4:
        lower = E.zona_coarse(c0, c1, c0, c1, ncoarse, c0 > 0, c1 < ncoarse);
        if (h0 < cb[c0]) E.add_bins(h0, cb[c0], lower);
        if (cb[c1] > h1) E.minus(lower, sum_bins(ix, h1, cb[c1]));
        break;
    case 5:
        lower = E.zona_coarse(c0 - 1, c1, c0 - 1, c1, ncoarse, c0 > 1, c1 < ncoarse);
        if (h0 > cb[c0 - 1]) E.minus(lower, sum_bins(ix, cb[c0 - 1], h0));
        if (cb[c1] > h1) E.minus(lower, sum_bins(ix, h1, cb[c1]));
        break;
    default:
        lower = sum_bins(ix, h0, h1);
    }
    *loads = E.loads;
    return lower;
}

} // namespace orc
