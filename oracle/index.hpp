// CPU oracle: restatement of ibis::relic / ibis::bin build and range
// evaluation (FastBit 2.0.3).  TEST INFRASTRUCTURE ONLY -- see wah.hpp.
#pragma once
#include "wah.hpp"
#include <vector>

namespace orc {

enum { OPN = 0, OPLT = 1, OPLE = 2 };
enum { COL_I32 = 0, COL_U32 = 1, COL_U16 = 2, COL_F32 = 3, COL_F64 = 4 };
// Plan choice of the reference: COST_IAPI is bord::column::evaluateRange
// (arrays registered through the in-memory API, src/bord.cpp:5825-6058),
// COST_PARTITION is column::evaluateRange (data partitions with index
// files, src/column.cpp:6029-6207).
enum { COST_IAPI = 0, COST_PARTITION = 1 };
enum { PLAN_INDEX = 1, PLAN_SCAN = 3 };

struct Index {
    int kind = 0;                      // 1 relic, 2 bin
    uint32_t nrows = 0;
    std::vector<double> vals;          // relic values / bin bounds
    std::vector<double> minval, maxval;
    std::vector<BV*> bits;             // null for empty bins
    bool offsets_built = false;        // the constructor filled offset64
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

// part::doScan over all rows: integer columns convert the bounds to the
// element type first (src/part.cpp:11130-11237), floats compare as doubles.
BV scan(const void* col, int type, uint32_t n, double lo, int lop, double hi, int rop);

// relic::estimateCost / bin::estimateCost (src/irelic.cpp:1857-1902,
// src/ibin.cpp:6531-6569) as the index built under `model` has them.
double estimate_cost(const Index& ix, int type, double lo, int lop, double hi, int rop, int model);
// The scan cost the index cost is compared with.
double scan_cost(const Index& ix, int type, int model);
// The reference's choice: PLAN_INDEX or PLAN_SCAN; *empty when the range
// misses the column's [min, max] and nothing is evaluated.
int auto_plan(const Index& ix, const void* col, int type, double lo, int lop, double hi, int rop,
              int model, bool* empty);
// bin::estimate(expr, lower, upper) -- src/ibin.cpp:6722-6829; for a relic
// (src/irelic.h:58-62) lower is evaluate's result and upper is cleared.
// Rows of lower are hits, rows of upper candidates; a cleared upper (size
// 0) stands for lower itself.
void estimate(const Index& ix, const void* col, int type, double lo, int lop, double hi, int rop,
              BV& lower, BV& upper);
// plan 0 (the reference's choice), PLAN_INDEX or PLAN_SCAN.
BV evaluate_plan(const Index& ix, const void* col, int type, double lo, int lop, double hi, int rop,
                 int plan, int model, int* used);

} // namespace orc
