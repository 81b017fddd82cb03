// Equal-weight bin boundaries of ibis::bin (host code).
//
// bin::setBoundaries with an "equal-weight" specification
// (src/ibin.cpp:2975-3332) calls scanAndPartition (src/ibin.cpp:2389-2582),
// which histograms the distinct values (index::mapValues,
// src/index.cpp:3411-3479), groups them with index::divideCounts
// (src/index.cpp:3648-4206) and places each boundary with
// util::compactValue (src/util.cpp:1227-1312); setBoundaries then sorts the
// boundaries, drops duplicates and appends DBL_MAX.  The steps below follow
// those functions one for one, including their integer arithmetic, because
// the boundaries decide which rows share a bitmap.
#include "../../include/fastbit_hip.h"
#include "index_host.h"
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

namespace fbh_ix {

namespace {

// A read-only section [b, e) of a count array (array_t(rhs, begin, end),
// src/array_t.cpp:155-178: an end before the begin means "to the end").
struct Counts {
    const uint32_t* p;
    uint32_t n;
    uint32_t operator[](uint32_t i) const { return p[i]; }
    Counts sub(uint32_t b, uint32_t e) const {
        if (b > n) b = n;
        if (e > n || e < b) e = n;
        return Counts{p + b, e - b};
    }
};

// index::divideCounts(bdry, cnt): bdry.size() groups; group k holds the
// values [bdry[k-1], bdry[k]).
void divide_counts(std::vector<uint32_t>& bdry, const Counts& cnt) {
    if (bdry.empty()) return;
    const uint32_t nb = (uint32_t)bdry.size();
    const uint32_t ncnt = cnt.n;
    uint32_t i, j, avg = 0, top = 0;
    if (nb * 3 / 2 >= ncnt) {
        bdry.resize(ncnt);
        for (i = 0; i < ncnt; ++i) bdry[i] = i + 1;
        return;
    }
    std::vector<uint32_t> weight(nb, 0);
    for (i = 0; i < ncnt; ++i) {
        avg += cnt[i];
        if (top < cnt[i]) top = cnt[i];
    }
    avg = (avg + (nb >> 1)) / nb;
    if (top < avg) {            // no isolated heavy values
        top = cnt[0];
        i = 1;
        j = 0;
        auto close_incl = [&]() {      // the bin takes value i too
            weight[j] = top + cnt[i];
            bdry[j] = i + 1;
            ++j;
            ++i;
            top = i < ncnt ? cnt[i] : 0;
        };
        auto close_excl = [&]() {      // the bin ends before value i
            weight[j] = top;
            bdry[j] = i;
            ++j;
            top = cnt[i];
        };
        while (i < ncnt && j < nb) {
            if (top + cnt[i] < avg) {
                top += cnt[i];
            } else if (top + cnt[i] == avg) {
                weight[j] = avg;
                bdry[j] = i + 1;
                ++j;
                ++i;
                top = i < ncnt ? cnt[i] : 0;
            } else if (j > 0 && weight[j - 1] > avg) {
                if (top > 0.9 * avg) close_excl();
                else if (top + cnt[i] < 1.2 * avg) close_incl();
                else if (top > 0.7 * avg) close_excl();
                else if (top + cnt[i] < 1.4 * avg) close_incl();
                else close_excl();
            } else if (top + cnt[i] < 1.1 * avg) {
                close_incl();
            } else if (top > 0.8 * avg) {
                close_excl();
            } else if (top + cnt[i] < 1.3 * avg) {
                close_incl();
            } else if (top > 0.6 * avg) {
                close_excl();
            } else {
                close_incl();
            }
            ++i;
        }
        if (top > 0) {          // values not yet in a bin
            if (j < nb) {
                weight[j] = top;
                bdry[j] = ncnt;
                ++j;
            } else {
                while (i < ncnt) { top += cnt[i]; ++i; }
                if (weight[j - 1] + top < (avg << 1)) {
                    weight[j - 1] += top;
                    bdry[j - 1] = ncnt;
                } else {
                    weight.push_back(top);
                    bdry.push_back(ncnt);
                    j = (uint32_t)bdry.size();
                }
            }
        }
        if (j < nb) {           // too many values in the first j bins
            bool dosplit = false;
            do {
                for (i = 1, top = 0; i < j; ++i)
                    if (weight[i] >= weight[top]) top = i;
                dosplit = false;
                for (i = top; i < j; ++i) {
                    if (i > 0) dosplit = bdry[i] > bdry[i - 1] + 1;
                    else dosplit = bdry[0] > 1;
                    if (dosplit) {
                        --bdry[i];
                        weight[i] -= cnt[bdry[i]];
                        if (i + 1 < j) {
                            weight[i + 1] += cnt[bdry[i]];
                        } else {
                            weight[i + 1] = cnt[bdry[i]];
                            bdry[i + 1] = ncnt;
                        }
                    }
                }
                j += dosplit ? 1u : 0u;
            } while (j < nb && dosplit);
            if (j < nb) {
                bdry.resize(j);
                weight.resize(j);
            }
        }
        // move borders to even out neighbouring bins
        bool doadjust = bdry.size() > 2;
        while (doadjust) {
            int diff = (int)(weight[1] - weight[0]);
            j = 1;
            for (i = 2; i < bdry.size(); ++i) {
                const int tmp = (int)(weight[i] - weight[i - 1]);
                if ((diff >= 0 ? diff : -diff) < (tmp >= 0 ? tmp : -tmp)) {
                    diff = tmp;
                    j = i;
                }
            }
            doadjust = false;
            if (diff > 0) {
                if (weight[j - 1] + cnt[bdry[j - 1]] < weight[j]) {
                    diff >>= 1;
                    doadjust = true;
                    if (cnt[bdry[j - 1]] > (uint32_t)diff) {
                        weight[j - 1] += cnt[bdry[j - 1]];
                        weight[j] -= cnt[bdry[j - 1]];
                        ++bdry[j - 1];
                    } else {
                        for (i = bdry[j - 1] + 1, top = cnt[bdry[j - 1]]; top <= (uint32_t)diff; ++i)
                            top += cnt[i];
                        --i;
                        top -= cnt[i];
                        weight[j - 1] += top;
                        weight[j] -= top;
                        bdry[j - 1] = i;
                    }
                } else if (j > 1 && weight[j - 1] + cnt[bdry[j - 1]] - cnt[bdry[j - 2]] < weight[j]) {
                    doadjust = true;
                    for (i = j - 1; doadjust && i > 1; --i) {
                        if (weight[i - 1] + cnt[bdry[i - 1]] < weight[j]) break;
                        doadjust = (weight[i - 1] + cnt[bdry[i - 1]] - cnt[bdry[i - 2]]) < weight[j];
                    }
                    if (i == 1 && doadjust) doadjust = weight[0] + cnt[bdry[0]] < weight[j];
                    if (doadjust) {
                        while (i <= j) {
                            weight[i - 1] += cnt[bdry[i - 1]];
                            weight[i] -= cnt[bdry[i - 1]];
                            ++bdry[i - 1];
                            ++i;
                        }
                    }
                }
            } else if (diff < 0) {
                if (weight[j - 1] > weight[j] + cnt[bdry[j - 1] - 1]) {
                    doadjust = true;
                    diff = -diff / 2;
                    if (cnt[bdry[j - 1] - 1] > (uint32_t)diff) {
                        --bdry[j - 1];
                        weight[j] += cnt[bdry[j - 1]];
                        weight[j - 1] -= cnt[bdry[j - 1]];
                    } else {
                        for (i = bdry[j - 1] - 2, top = cnt[bdry[j - 1] - 1]; top + cnt[i] <= (uint32_t)diff; --i)
                            top += cnt[i];
                        ++i;
                        bdry[j - 1] = i;
                        weight[j] += top;
                        weight[j - 1] -= top;
                    }
                } else if (weight[j - 1] - cnt[bdry[j - 1] - 1] > weight[j] - cnt[bdry[j] - 1]) {
                    doadjust = j + 1 < weight.size();
                    for (i = j + 1; doadjust && i < weight.size(); ++i) {
                        if (weight[j - 1] > weight[i] + cnt[bdry[i - 1] - 1]) break;
                        doadjust = (i + 1 < weight.size()) &&
                                   (weight[i - 1] - cnt[bdry[i - 1] - 1] > weight[i] - cnt[bdry[i] - 1]);
                    }
                    if (doadjust) {
                        while (i >= j) {
                            --bdry[i - 1];
                            weight[i] += cnt[bdry[i - 1]];
                            weight[i - 1] -= cnt[bdry[i - 1]];
                            --i;
                        }
                    }
                }
            }
        }
        return;
    }
    // values with very large counts get bins of their own
    for (i = 0, j = 0; i < ncnt && j < nb; ++i) {
        if (cnt[i] >= avg) {
            weight[j] = i;      // positions of the heavy values
            ++j;
        }
    }
    if (i < ncnt || j >= nb) {  // all values have about equal counts
        avg = ncnt / nb;
        j = ncnt % nb;
        for (i = 0, top = 0; i < j; ++i) {
            top += avg + 1;
            bdry[i] = top;
        }
        for (; i < nb; ++i) {
            top += avg;
            bdry[i] = top;
        }
        return;
    }
    weight.resize(j);
    std::vector<uint32_t> cnt2(j + 1);
    cnt2[0] = 0;
    avg = 0;
    for (i = 0; i < weight[0]; ++i) cnt2[0] += cnt[i];
    avg += cnt2[0];
    for (i = 1; i < j; ++i) {
        cnt2[i] = 0;
        for (uint32_t ki = weight[i - 1] + 1; ki < weight[i]; ++ki) cnt2[i] += cnt[ki];
        avg += cnt2[i];
    }
    cnt2[j] = 0;
    for (i = weight.back() + 1; i < ncnt; ++i) cnt2[j] += cnt[i];
    avg += cnt2[j];
    avg = (avg > nb - j) ? (avg + ((nb - j) >> 1)) / (nb - j) : 1;
    top = avg >> 1;
    std::vector<uint32_t> nb2(j + 1);
    for (i = 0; i <= j; ++i) {
        nb2[i] = (top + cnt2[i]) / avg;
        if (nb2[i] == 0 && cnt2[i] > 0) {
            nb2[i] = 1;
        } else if (i == j) {
            if (nb2[i] > ncnt - weight.back() - 1) nb2[i] = ncnt - weight.back() - 1;
        } else if (i > 0) {
            if (nb2[i] > weight[i] - weight[i - 1] - 1) nb2[i] = weight[i] - weight[i - 1] - 1;
        } else if (i == 0 && nb2[0] > weight[0]) {
            nb2[0] = weight[0];
        }
    }
    for (i = 0, avg = j; i < j + 1; ++i) avg += nb2[i];
    while (avg > nb) {          // fewer bins: take one from the interval with the most
        top = 0;
        double frac = DBL_MAX;
        if (nb2[0] > 1) frac = (double)cnt2[0] / nb2[0];
        for (i = 1; i <= j; ++i) {
            if (nb2[i] > 1) {
                if (frac < DBL_MAX) {
                    if (frac * nb2[i] < cnt2[i]) {
                        top = i;
                        frac = (double)cnt2[i] / nb2[i];
                    } else if (frac * nb2[i] == cnt2[i] && cnt2[i] > cnt2[top]) {
                        top = i;
                        frac = (double)cnt2[i] / nb2[i];
                    }
                } else {
                    top = i;
                    frac = (double)cnt2[i] / nb2[i];
                }
            }
        }
        if (frac == DBL_MAX) break;
        --nb2[top];
        --avg;
    }
    while (avg < nb) {          // more bins: give one to the most crowded interval
        top = 0;
        double frac = nb2[0] < weight[0] ? (nb2[0] > 0 ? (double)(cnt2[0] / nb2[0]) : (double)cnt2[0]) : 0.0;
        for (i = 1; i <= j; ++i) {
            // the reference's bound here compares against a bool for inner
            // intervals (src/index.cpp:4063-4064); kept as written
            const uint32_t lim = i < j ? (uint32_t)(nb2[i] > weight[i] - weight[i - 1] - 1)
                                       : ncnt - weight.back() - 1;
            if (nb2[i] > 0 && nb2[i] < lim) {
                if (frac * nb2[i] > cnt2[i]) {
                    top = i;
                    frac = (double)cnt2[i] / nb2[i];
                } else if (frac * nb2[i] == cnt2[i] && cnt2[i] > cnt2[top]) {
                    top = i;
                    frac = (double)cnt2[i] / nb2[i];
                } else if (frac <= 0.0) {
                    top = i;
                    frac = (double)cnt2[i] / nb2[i];
                }
            }
        }
        if (frac == 0.0) break;
        ++nb2[top];
        ++avg;
    }
    // establish the borders
    if (nb2[0] > 1) {
        bdry.assign(nb2[0], 0);
        divide_counts(bdry, cnt.sub(0, weight[0]));
    } else if (nb2[0] == 1) {
        bdry.assign(1, weight[0]);
    } else {
        bdry.clear();
    }
    for (i = 0; i < j; ++i) {
        const uint32_t off = weight[i] + 1;
        bdry.push_back(off);
        if (nb2[i + 1] > 1) {
            std::vector<uint32_t> bnd(nb2[i + 1], 0);
            divide_counts(bnd, cnt.sub(off, i + 1 < j ? weight[i + 1] : ncnt));
            for (uint32_t x : bnd) bdry.push_back(off + x);
        } else if (nb2[i + 1] == 1) {
            bdry.push_back(i + 1 < j ? weight[i + 1] : ncnt);
        }
    }
}

} // namespace

void equal_weight_bounds(const double* val, const uint32_t* cnt, uint32_t ncnt, uint32_t nbins,
                         bool is_float, double col_upper, std::vector<double>& bounds) {
    bounds.clear();
    if (ncnt > nbins * 3 / 2) {
        std::vector<uint32_t> bnds(nbins, 0);
        divide_counts(bnds, Counts{cnt, ncnt});
        if (is_float) {
            // src/ibin.cpp:2495-2527
            if (val[0] >= 0) bounds.push_back(0.0);
            else bounds.push_back(compact_value(val[0], -DBL_MAX));
            for (uint32_t b : bnds) {
                double tmp;
                if (b == 1) {
                    tmp = compact_value(0.5 * (val[0] + val[1]), val[1]);
                } else if (b < ncnt) {
                    tmp = compact_value(val[b - 1], val[b]);
                } else {
                    tmp = col_upper;
                    if (val[ncnt - 1] <= tmp) tmp = compact_value(val[ncnt - 1], tmp);
                    else tmp = compact_value(val[ncnt - 1], DBL_MAX);
                }
                bounds.push_back(tmp);
            }
        } else {
            // src/ibin.cpp:2529-2551: a heavy value gets a bin of its own
            uint32_t avg = 0;
            for (uint32_t i = 0; i < ncnt; ++i) avg += cnt[i];
            avg /= nbins;
            bool skip = false;
            for (size_t k = 0; k < bnds.size() && bnds[k] < ncnt; ++k) {
                const uint32_t b = bnds[k];
                if (skip) {
                    skip = false;
                } else {
                    bounds.push_back(val[b]);
                    if (cnt[b] > avg) {
                        const uint32_t next = b + 1;
                        if (next < ncnt && val[b] + 1 < val[next]) {
                            bounds.push_back(val[b] + 1);
                            skip = true;
                        }
                    }
                }
            }
        }
    } else if (ncnt > 1) {
        for (uint32_t i = 0; i < ncnt; ++i) bounds.push_back(val[i]);
    } else if (ncnt > 0) {
        if (std::fabs(val[0] - 1) < 0.5) {
            bounds.push_back(0.0);
            bounds.push_back(2.0);
        } else {
            bounds.push_back(compact_value(val[0], -DBL_MAX));
            bounds.push_back(compact_value(val[0], DBL_MAX));
        }
    }
    // setBoundaries: sorted, unique, DBL_MAX last (src/ibin.cpp:3275-3315)
    if (bounds.size() > 1) {
        std::sort(bounds.begin(), bounds.end());
        size_t i = 0;
        for (size_t k = 1; k < bounds.size(); ++k) {
            if (bounds[k] > bounds[i]) {
                if (k > i + 1) bounds[i + 1] = bounds[k];
                ++i;
            }
        }
        bounds.resize(i + 1);
    }
    if (!bounds.empty()) bounds.push_back(DBL_MAX);
}

} // namespace fbh_ix

extern "C" int fbh_bin_equal_weight(const double* vals, const uint32_t* counts, uint32_t ncnt,
                                    uint32_t nbins, int is_float, double col_upper, double* bounds,
                                    uint32_t cap, uint32_t* nbounds) {
    if ((ncnt && (!vals || !counts)) || !nbounds || nbins == 0) return FBH_ERR_ARG;
    std::vector<double> b;
    fbh_ix::equal_weight_bounds(vals, counts, ncnt, nbins, is_float != 0, col_upper, b);
    *nbounds = (uint32_t)b.size();
    if (bounds) std::memcpy(bounds, b.data(), sizeof(double) * std::min<size_t>(cap, b.size()));
    return FBH_OK;
}
