// Host-side restatements of the reference's small sequential routines that
// sit around the index kernels: relic::locate, bin::locate (value and
// range forms), util::compactValue, and the bound adjustment that
// part::doScan applies before comparing integer columns.  They run on a few
// doubles per query; everything proportional to rows or words runs on the
// GPU.
#pragma once
#include <stdint.h>
#include <vector>
#include <cmath>
#include <cfloat>
#include <limits>

namespace fbh_ix {

enum { OP_NONE = 0, OP_LT = 1, OP_LE = 2 };

// ibis::relic::locate(double) -- src/irelic.cpp:1362-1411
inline uint32_t relic_locate(const std::vector<double>& vals, double val) {
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

// ibis::relic::locate(qContinuousRange, hit0, hit1) for the operator
// combinations lo {<,<=} x {<,<=} hi and one-sided forms
// (src/irelic.cpp:1416-1809, cases OP_LT / OP_LE on the left).
inline void relic_locate_range(const std::vector<double>& vals, double lo, int lop, double hi,
                               int rop, uint32_t& hit0, uint32_t& hit1) {
    const uint32_t nval = (uint32_t)vals.size();
    const uint32_t bin0 = lop != OP_NONE ? relic_locate(vals, lo) : 0;
    const uint32_t bin1 = rop != OP_NONE ? relic_locate(vals, hi) : 0;
    if (lop == OP_LT) hit0 = bin0;
    else if (lop == OP_LE) hit0 = bin0 > 0 ? bin0 - (lo == vals[bin0 - 1]) : 0;
    else hit0 = 0;
    if (rop == OP_LT) hit1 = bin1 > 0 ? bin1 - (hi == vals[bin1 - 1]) : 0;
    else if (rop == OP_LE) hit1 = bin1;
    else hit1 = nval;
}

// ibis::bin::locate(double) -- src/ibin.cpp:673-738
inline uint32_t bin_locate(const std::vector<double>& bounds, double val) {
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

// ibis::bin::locate(expr, cand0, cand1, hit0, hit1) -- src/ibin.cpp:7596-8460
// for the left operators {undefined, <, <=} and right operators
// {undefined, <, <=} (LE/LT cases at 7854-7894 and their mirrors).
inline void bin_locate_range(const std::vector<double>& bounds, const std::vector<double>& minval,
                             const std::vector<double>& maxval, double lo, int lop, double hi,
                             int rop, uint32_t& cand0, uint32_t& cand1, uint32_t& hit0,
                             uint32_t& hit1) {
    const uint32_t nobs = (uint32_t)bounds.size();
    const uint32_t bin0 = lop != OP_NONE ? bin_locate(bounds, lo) : 0;
    const uint32_t bin1 = rop != OP_NONE ? bin_locate(bounds, hi) : 0;
    // left side
    if (lop == OP_LE) {
        if (bin0 >= nobs) { hit0 = nobs; cand0 = nobs; }
        else if (lo > maxval[bin0]) { hit0 = bin0 + 1; cand0 = bin0 + 1; }
        else if (lo <= minval[bin0]) { hit0 = bin0; cand0 = bin0; }
        else { hit0 = bin0 + 1; cand0 = bin0; }
    } else if (lop == OP_LT) {
        if (bin0 >= nobs) { hit0 = nobs; cand0 = nobs; }
        else if (lo >= maxval[bin0]) { hit0 = bin0 + 1; cand0 = bin0 + 1; }
        else if (lo < minval[bin0]) { hit0 = bin0; cand0 = bin0; }
        else { hit0 = bin0 + 1; cand0 = bin0; }
    } else {
        hit0 = 0; cand0 = 0;
    }
    // right side
    if (rop == OP_LT) {
        if (bin1 >= nobs) { hit1 = nobs; cand1 = nobs; }
        else if (hi > maxval[bin1]) { hit1 = bin1 + 1; cand1 = bin1 + 1; }
        else if (hi <= minval[bin1]) { hit1 = bin1; cand1 = bin1; }
        else { hit1 = bin1; cand1 = bin1 + 1; }
    } else if (rop == OP_LE) {
        if (bin1 >= nobs) { hit1 = nobs; cand1 = nobs; }
        else if (hi >= maxval[bin1]) { hit1 = bin1 + 1; cand1 = bin1 + 1; }
        else if (hi < minval[bin1]) { hit1 = bin1; cand1 = bin1; }
        else { hit1 = bin1; cand1 = bin1 + 1; }
    } else {
        hit1 = nobs; cand1 = nobs;
    }
}

// ibis::util::compactValue -- src/util.cpp:1227-1312
inline double compact_value(double left, double right, double start = 0.0) {
    if (left == right) return left;
    if (!(left != right)) return 0.0;
    if (left > right) { double t = left; left = right; right = t; }
    if (left < 0 && right >= 0) return 0.0;
    if (left < 1.0 && right >= 1.0) return 1.0;
    if (left < -1.0 && right >= -1.0) return -1.0;
    double diff, sep;
    if (left == 0.0) {
        diff = std::floor(std::log10(right));
        sep = std::pow(10.0, diff);
        if (sep > right) {
            if (diff >= -3.0 && diff < 3.0) sep *= 0.5;
            else sep *= 0.1;
        }
    } else if (right < 0.0 && right * 10 > left) {
        diff = std::ceil(std::log10(-right));
        sep = -std::pow(10.0, diff);
        if (sep > right) {
            if (diff >= -3.0 && diff <= 3.0) sep += sep;
            else sep *= 10;
        }
    } else if (left > 0.0 && right > 10 * left) {
        diff = std::ceil(std::log10(left));
        sep = std::pow(10.0, diff);
        if (sep <= left) {
            if (sep >= -3.0 && diff <= 3.0) sep += sep;
            else sep *= 10.0;
        }
    } else {
        diff = std::pow(10.0, std::ceil(FLT_EPSILON + std::log10(right - left)));
        if (!(start > left && start <= right)) start = 0.5 * (right + left);
        sep = std::floor(0.5 + start / diff) * diff;
        if (!(sep > left && sep <= right)) {
            diff /= 2;
            sep = std::floor(0.5 + start / diff) * diff;
            if (!(sep > left && sep <= right)) {
                diff /= 5;
                sep = std::floor(0.5 + start / diff) * diff;
                if (!(sep > left && sep <= right)) {
                    diff /= 2;
                    sep = std::floor(0.5 + start / diff) * diff;
                    if (!(sep > left && sep <= right)) {
                        diff /= 2;
                        sep = std::floor(0.5 + start / diff) * diff;
                    }
                }
            }
        }
    }
    if (!(sep > left && sep <= right)) sep = right;
    return sep;
}

// Equal-weight bin boundaries from the sorted distinct values and their
// counts (bin::scanAndPartition + setBoundaries; bin_bounds.cpp).
void equal_weight_bounds(const double* val, const uint32_t* cnt, uint32_t ncnt, uint32_t nbins,
                         bool is_float, double col_upper, std::vector<double>& bounds);

// part::doScan's bound adjustment for an integer column of type T
// (src/part.cpp:11130-11237): bounds are cast to T and the operator moves
// between < and <= when the cast changed the value.
template <class T>
inline void int_bounds(double lo, int& lop, double hi, int& rop, int64_t& ilo, int64_t& ihi) {
    const double tmin = (double)std::numeric_limits<T>::min();
    const double tmax = (double)std::numeric_limits<T>::max();
    T lb = 0, rb = 0;
    if (lop == OP_LT) {
        if (lo < tmin) { lb = std::numeric_limits<T>::min(); lop = OP_LE; }
        else if (lo > tmax) { lb = std::numeric_limits<T>::max(); }
        else lb = (T)lo;
    } else if (lop == OP_LE) {
        if (lo < tmin) { lb = std::numeric_limits<T>::min(); lop = OP_LE; }
        else if (lo > tmax) { lb = std::numeric_limits<T>::max(); lop = OP_LT; }
        else { lb = (T)lo; if ((double)lb != lo) lop = OP_LT; }
    }
    if (rop == OP_LE) {
        if (hi < tmin) { rb = std::numeric_limits<T>::min(); rop = OP_LT; }
        else if (hi > tmax) { rb = std::numeric_limits<T>::max(); }
        else rb = (T)hi;
    } else if (rop == OP_LT) {
        if (hi < tmin) { rb = std::numeric_limits<T>::min(); }
        else if (hi > tmax) { rb = std::numeric_limits<T>::max(); rop = OP_LE; }
        else { rb = (T)hi; if ((double)rb < hi) rop = OP_LE; }
    }
    ilo = (int64_t)lb;
    ihi = (int64_t)rb;
}

} // namespace fbh_ix
