// CPU driver of the WAH step core (wah_core.h), used by the CPU test suite to
// check the partitioned algorithm that the gfx950 kernels run: the position
// space is cut at caller-chosen points, each piece is advanced to its first
// synchronisation point, walked, summarized and written exactly as one GPU
// thread does.  Built as a host-only library (libfbsim.so); not part of the product
// entry points.
#include "wah_core.h"
#include <vector>
#include <cstring>
#include <algorithm>

using namespace wah;

namespace {
struct VecWriter {
    std::vector<word_t>* out;
    std::vector<uint32_t>* heads;
    void operator()(uint32_t slot, word_t w, uint32_t headpos) {
        if (slot >= out->size()) { out->resize(slot + 1); heads->resize(slot + 1); }
        (*out)[slot] = w;
        (*heads)[slot] = headpos;
    }
};
}

extern "C" {

// a, b: code words only (no active word), both covering `groups` 31-bit
// groups.  splits: nominal cut positions (any order, any value in [0,groups]).
// mode: 0 the compressed *_c2 result, MODE_CANON the canonical result,
// MODE_RAW the *_c2 loop's own words.  Writes the result words to out
// (capacity cap) and their start positions to heads; returns the word count
// or a negative value.
long fbsim_op(const uint32_t* a, uint32_t na, const uint32_t* b, uint32_t nb,
              uint32_t groups, int op, int mode, const uint32_t* splits,
              uint32_t nsplits, uint32_t* out, uint32_t* heads, uint32_t cap) {
    ArrayReader ra{a}, rb{b};
    // start positions of every code word (plus sentinel)
    std::vector<uint32_t> sa(na + 1), sb(nb + 1);
    sa[0] = 0; for (uint32_t i = 0; i < na; ++i) sa[i + 1] = sa[i] + word_len(a[i]);
    sb[0] = 0; for (uint32_t i = 0; i < nb; ++i) sb[i + 1] = sb[i] + word_len(b[i]);
    if (sa[na] != groups || sb[nb] != groups) return -1;
    std::vector<uint32_t> cuts(splits, splits + nsplits);
    cuts.push_back(0);
    std::vector<uint32_t> starts;   // sync points (positions)
    std::vector<uint32_t> ias, ibs, pas, pbs;
    for (uint32_t c : cuts) {
        if (c > groups) c = groups;
        // code word containing c in each operand, then advance to common
        uint32_t ia = (uint32_t)(std::upper_bound(sa.begin(), sa.end(), c) - sa.begin()) - 1;
        uint32_t ib = (uint32_t)(std::upper_bound(sb.begin(), sb.end(), c) - sb.begin()) - 1;
        if (ia > na) ia = na;
        if (ib > nb) ib = nb;
        uint32_t pa = sa[ia], pb = sb[ib];
        if (pa < c && ia < na) { pa = sa[ia + 1]; ++ia; }
        if (pb < c && ib < nb) { pb = sb[ib + 1]; ++ib; }
        uint32_t p = find_sync(&ra, &rb, ia, pa, ib, pb, groups, (mode & MODE_RAW) != 0);
        starts.push_back(p); ias.push_back(ia); ibs.push_back(ib);
        pas.push_back(pa); pbs.push_back(pb);
    }
    // sort and dedupe the sync points
    std::vector<size_t> ord(starts.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return starts[x] < starts[y]; });
    std::vector<uint32_t> P, IA, IB, PA, PB;
    for (size_t k : ord) {
        if (!P.empty() && P.back() == starts[k]) continue;
        P.push_back(starts[k]); IA.push_back(ias[k]); IB.push_back(ibs[k]);
        PA.push_back(pas[k]); PB.push_back(pbs[k]);
    }
    if (P.back() != groups) { P.push_back(groups); IA.push_back(na); IB.push_back(nb); PA.push_back(groups); PB.push_back(groups); }
    const size_t nr = P.size() - 1;
    // count pass
    std::vector<Summary> sum(nr);
    for (size_t r = 0; r < nr; ++r) {
        CountSink cs; cs.init();
        walk(cs, &ra, &rb, IA[r], PA[r], IB[r], PB[r], P[r], P[r + 1], op, (uint32_t)mode);
        sum[r] = cs.s;
    }
    // exclusive prefixes
    std::vector<Summary> pre(nr);
    Summary acc = summary_empty();
    for (size_t r = 0; r < nr; ++r) { pre[r] = acc; acc = combine(acc, sum[r]); }
    // write pass
    std::vector<word_t> res;
    std::vector<uint32_t> hd;
    VecWriter vw{&res, &hd};
    for (size_t r = 0; r < nr; ++r) {
        WriteSink<VecWriter> ws;
        ws.init(&vw, pre[r], P[r]);
        walk(ws, &ra, &rb, IA[r], PA[r], IB[r], PB[r], P[r], P[r + 1], op, (uint32_t)mode);
        if (r + 1 == nr) ws.finish();
    }
    if (res.size() != acc.count) return -2;
    if (res.size() > cap) return -3;
    std::memcpy(out, res.data(), res.size() * sizeof(word_t));
    std::memcpy(heads, hd.data(), hd.size() * sizeof(uint32_t));
    return (long)res.size();
}

// The *_c2 result as the group-domain kernels write it: the canonical words
// of the op's group values, with every single-group run of all-0 or all-1
// groups written as a count-1 fill where c2_rep_at finds a fill token
// there.  Returns the word count or a negative value.
long fbsim_c2_canon_rep(const uint32_t* a, uint32_t na, const uint32_t* b, uint32_t nb, uint32_t groups,
                        int op, uint32_t* out, uint32_t cap) {
    ArrayReader ra{a}, rb{b};
    std::vector<uint32_t> sa(na + 1), sb(nb + 1);
    sa[0] = 0; for (uint32_t i = 0; i < na; ++i) sa[i + 1] = sa[i] + word_len(a[i]);
    sb[0] = 0; for (uint32_t i = 0; i < nb; ++i) sb[i + 1] = sb[i] + word_len(b[i]);
    if (sa[na] != groups || sb[nb] != groups) return -1;
    std::vector<word_t> v(groups);
    for (uint32_t i = 0, g = 0; i < na; ++i)
        for (uint32_t k = 0; k < word_len(a[i]); ++k) v[g++] = group_value(a[i]);
    for (uint32_t i = 0, g = 0; i < nb; ++i)
        for (uint32_t k = 0; k < word_len(b[i]); ++k, ++g) v[g] = apply_op(v[g], group_value(b[i]), op);
    std::vector<word_t> res;
    for (uint32_t g = 0; g < groups;) {
        const uint32_t c = class_of_value(v[g]);
        uint32_t e = g + 1;
        if (c != C_MIXED)
            while (e < groups && v[e] == v[g]) ++e;
        if (c == C_MIXED) {
            res.push_back(v[g]);
        } else if (e - g >= 2) {
            res.push_back((c == C_ONE ? HEADER1 : HEADER0) | (e - g));
        } else {
            const uint32_t ia = (uint32_t)(std::upper_bound(sa.begin(), sa.end(), g) - sa.begin()) - 1;
            const uint32_t ib = (uint32_t)(std::upper_bound(sb.begin(), sb.end(), g) - sb.begin()) - 1;
            const bool fill = c2_rep_at(&ra, &rb, ia, sa[ia], ib, sb[ib], g, op);
            res.push_back(fill ? ((c == C_ONE ? HEADER1 : HEADER0) | 1u) : v[g]);
        }
        g = e;
    }
    if (res.size() > cap) return -3;
    std::memcpy(out, res.data(), res.size() * sizeof(word_t));
    return (long)res.size();
}

} // extern "C"
