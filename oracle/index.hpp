// CPU oracle: restatement of ibis::relic / ibis::bin build and range
// evaluation (FastBit 2.0.3).  TEST INFRASTRUCTURE ONLY -- see wah.hpp.
#pragma once
#include "wah.hpp"
#include <vector>

namespace orc {

enum { OPN = 0, OPLT = 1, OPLE = 2 };
enum { COL_I32 = 0, COL_U32 = 1, COL_U16 = 2, COL_F32 = 3, COL_F64 = 4 };

struct Index {
    int kind = 0;                      // 1 relic, 2 bin
    uint32_t nrows = 0;
    std::vector<double> vals;          // relic values / bin bounds
    std::vector<double> minval, maxval;
    std::vector<BV*> bits;             // null for empty bins
    ~Index() { for (auto b : bits) delete b; }
};

// relic(col, card, ind) -- src/irelic.cpp:160-211
Index* relic_card(const uint32_t* ind, uint32_t n, uint32_t card);
// index::mapValues + relic::construct -- src/index.cpp:2552-2611,
// src/irelic.cpp:878-925 (column values given as doubles)
Index* relic_values(const double* v, uint32_t n);
// bin::binning with given bounds -- src/ibin.cpp:3350-3449
Index* bin_bounds(const double* v, uint32_t n, const double* bounds, uint32_t nb);

// relic::write / bin::write -- src/irelic.cpp:644-667, src/ibin.cpp:5836-5860
void write_index(const Index& ix, std::vector<double>& keys, std::vector<int64_t>& offs,
                 std::vector<word_t>& bms);

// relic::evaluate / bin::evaluate -- src/irelic.cpp:1827-1839,
// src/ibin.cpp:6614-6713; the edge check follows part::doScan's
// comparison rules for the column type (src/part.cpp:11130-12219).
// Returns the hit bitvector.
BV evaluate(const Index& ix, const void* col, int type, double lo, int lop, double hi, int rop);

} // namespace orc
