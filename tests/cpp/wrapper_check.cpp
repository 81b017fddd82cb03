// Self-checking use of include/fastbit_hip.hpp: bitvector operators, index
// evaluation, attach, sumBins and an expression batch, each compared with
// brute force over plain bit arrays on the host.  Exit status 0 when every
// check holds.
#include "fastbit_hip.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

int failures = 0;
#define EXPECT(c)                                                      \
    do {                                                               \
        if (!(c)) { std::printf("FAILED %s:%d %s\n", __FILE__, __LINE__, #c); ++failures; } \
    } while (0)

uint64_t rng_state = 0x9E3779B97F4A7C15ull;
uint32_t next() {
    rng_state ^= rng_state << 13; rng_state ^= rng_state >> 7; rng_state ^= rng_state << 17;
    return (uint32_t)(rng_state >> 11);
}

// bits with runs of ones and zeros (so results compress into fills)
std::vector<uint8_t> runs(uint32_t n, uint32_t mean) {
    std::vector<uint8_t> b(n);
    uint8_t v = next() & 1;
    for (uint32_t i = 0; i < n;) {
        uint32_t len = 1 + next() % (2 * mean);
        for (uint32_t k = 0; k < len && i < n; ++k, ++i) b[i] = v && (mean > 40 || (next() & 3));
        v ^= 1;
    }
    return b;
}

// the serialized form of an uncompressed bitvector: one literal per 31
// bits (bit 30 is the first), then the active word and its bit count
std::vector<uint32_t> literal_serial(const std::vector<uint8_t>& b) {
    const uint32_t n = (uint32_t)b.size(), g = n / 31, anb = n % 31;
    std::vector<uint32_t> w;
    for (uint32_t i = 0; i < g; ++i) {
        uint32_t x = 0;
        for (uint32_t j = 0; j < 31; ++j) x = (x << 1) | b[31 * i + j];
        w.push_back(x);
    }
    if (anb) {
        uint32_t x = 0;
        for (uint32_t j = 0; j < anb; ++j) x = (x << 1) | b[31 * g + j];
        w.push_back(x);
    }
    w.push_back(anb);
    return w;
}

std::vector<uint32_t> ones_of(const std::vector<uint8_t>& b) {
    std::vector<uint32_t> r;
    for (uint32_t i = 0; i < b.size(); ++i) if (b[i]) r.push_back(i);
    return r;
}

void check_ops(uint32_t n) {
    const auto x = runs(n, 1 + next() % 200), y = runs(n, 1 + next() % 200);
    const fbh::bitvector a(literal_serial(x)), b(literal_serial(y));
    for (int op = 0; op < 4; ++op) {
        std::vector<uint8_t> e(n);
        for (uint32_t i = 0; i < n; ++i)
            e[i] = op == 0 ? (x[i] & y[i]) : op == 1 ? (x[i] | y[i]) : op == 2 ? (x[i] ^ y[i]) : (x[i] & !y[i]);
        fbh::bitvector c(a);
        if (op == 0) c &= b; else if (op == 1) c |= b; else if (op == 2) c ^= b; else c -= b;
        const fbh::bitvector d = op == 0 ? a & b : op == 1 ? a | b : op == 2 ? a ^ b : a - b;
        const auto want = ones_of(e);
        EXPECT(c.size() == n && d.size() == n);
        EXPECT(c.cnt() == want.size() && d.cnt() == want.size());
        EXPECT(c.rids() == want && d.rids() == want);
        // the out-of-place result is the routine's words, not compressed:
        // compressed, it is the in-place result
        std::vector<uint32_t> sc, sd;
        c.write(sc);
        fbh::bitvector dc(d);
        dc.compress();
        dc.write(sd);
        EXPECT(sc == sd);
        c.flip();
        EXPECT(c.cnt() == n - want.size());
    }
}

void check_index() {
    const uint32_t n = 100003;
    std::vector<int32_t> col(n);
    for (auto& v : col) v = (int32_t)(next() % 50) - 10;
    auto count = [&](double lo, double hi) {       // lo <= x < hi
        uint64_t c = 0;
        for (int32_t v : col) c += (lo <= v && v < hi);
        return c;
    };
    fbh::index ix = fbh::index::relic(col.data(), FBH_COL_INT32, n);
    EXPECT(ix.getNRows() == n && ix.numBitvectors() == 50);
    fbh::index at = fbh::index::attach(col.data(), FBH_COL_INT32, n, ix.deconstruct());
    for (int q = 0; q < 12; ++q) {
        const double lo = (double)((int)(next() % 60) - 15), hi = lo + (double)(next() % 30);
        const uint64_t want = count(lo, hi);
        for (int plan : {FBH_PLAN_AUTO, FBH_PLAN_LITERAL, FBH_PLAN_COMPRESSED, FBH_PLAN_SCAN}) {
            EXPECT(ix.evaluate(lo, FBH_OP_LE, hi, FBH_OP_LT, plan) == want);
            EXPECT(at.evaluate(lo, FBH_OP_LE, hi, FBH_OP_LT, plan) == want);
        }
        EXPECT(ix.rids(lo, FBH_OP_LE, hi, FBH_OP_LT).size() == want);
        EXPECT(ix.evaluate_bits(lo, FBH_OP_LE, hi, FBH_OP_LT).cnt() == want);
    }
    EXPECT(ix.sumBins(0, 50) == n);
    EXPECT(ix.sumBins(10, 20) == count(0, 10));
    // two-level forms over the same column (ibis::fuzz, ibis::zona)
    {
        fbh::index fz = fbh::index::fuzz(col.data(), FBH_COL_INT32, n, 6);
        fbh::index zn = fbh::index::zona(col.data(), FBH_COL_INT32, n);
        for (int q = 0; q < 8; ++q) {
            const double lo = (double)((int)(next() % 60) - 15), hi = lo + (double)(next() % 40);
            const uint64_t want = count(lo, hi);
            EXPECT(fz.evaluate(lo, FBH_OP_LE, hi, FBH_OP_LT, FBH_PLAN_LITERAL) == want);
            EXPECT(zn.evaluate(lo, FBH_OP_LE, hi, FBH_OP_LT, FBH_PLAN_AUTO) == want);
        }
    }
    uint32_t c0, h0, h1, c1;
    ix.locate(0, FBH_OP_LE, 10, FBH_OP_LT, c0, h0, h1, c1);
    EXPECT(h0 == 10 && h1 == 20);
    // (x in [0, 5) OR x in [20, 30)) AND NOT x <= 25
    std::vector<fbh_term> t = {{ix.handle(), 0, FBH_OP_LE, 5, FBH_OP_LT},
                               {ix.handle(), 20, FBH_OP_LE, 30, FBH_OP_LT},
                               {ix.handle(), 0, FBH_OP_NONE, 25, FBH_OP_LE}};
    fbh::batch e = fbh::batch::expressions(t, {0, 3}, {0, 1, FBH_EXPR_OR, 2, FBH_EXPR_NOT, FBH_EXPR_AND}, {0, 6});
    EXPECT(e.run()[0] == count(26, 30));
    std::vector<uint64_t> off;
    std::vector<uint32_t> rids;
    e.hit_rows(off, rids);
    EXPECT(off.size() == 2 && off[1] == count(26, 30) && rids.size() == off[1]);
    bool rows_ok = true;
    for (size_t i = 0; i < rids.size(); ++i)
        rows_ok = rows_ok && col[rids[i]] >= 26 && col[rids[i]] < 30 && (i == 0 || rids[i] > rids[i - 1]);
    EXPECT(rows_ok);
    EXPECT(e.hit_vector(0).cnt() == count(26, 30));
    fbh::batch cj = fbh::batch::conjunctions(t, {0, 2});
    EXPECT(cj.run()[0] == 0);
    bool threw = false;
    try {
        fbh::batch::expressions(t, {0, 3}, {0, FBH_EXPR_AND}, {0, 2});
    } catch (const fbh::error& err) {
        threw = err.code == FBH_ERR_ARG;
    }
    EXPECT(threw);
}

} // namespace

int main() {
    for (uint32_t n : {1u, 31u, 62u, 1000u, 100003u, 3000017u}) check_ops(n);
    check_index();
    std::printf(failures ? "wrapper checks failed: %d\n" : "wrapper ok\n", failures);
    return failures ? 1 : 0;
}
