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

// ---------------------------------------------------------------------------
// Two-level encodings over a relic: ibis::fuzz (interval-equality coarse
// level) and ibis::zona (equality-equality), src/irelic.h:540-712.
enum { ENC_FUZZ = 1, ENC_ZONA = 2 };
// What an evaluation did (the option numbers of fuzz::evaluate /
// zona::evaluate, 1..5; the other branches numbered here).
enum { TL_EMPTY = 0, TL_FINE = 1, TL_ALL = 6, TL_EQUAL = 7, TL_SAME_BIN = 8 };

struct TwoLevel {
    int enc = 0;
    std::vector<uint32_t> cbounds;     // coarse bin c holds fine bitmaps [cbounds[c], cbounds[c+1])
    std::vector<BV*> cbits;            // fuzz: ncoarse - ceil(ncoarse/2) + 1 bitmaps, zona: ncoarse
    std::vector<int64_t> off;          // offset64 of the fine level (bytes)
    std::vector<int64_t> coff;         // coffset64 (bytes)
    ~TwoLevel() { for (auto b : cbits) delete b; }
};

// fuzz::coarsen / zona::coarsen (src/ixfuzz.cpp:161-263, src/ixzona.cpp:
// 124-182) on the index built in memory: offset64 from bitvector::bytes(),
// coffset64 from the compressed coarse bitmaps' bytes().  ncoarse is the
// "ncoarse=" value of the index spec (fuzz only; values below 5 select the
// default rule).  Returns a level with no coarse bins where coarsen
// declines to build one.
TwoLevel* coarsen(const Index& ix, int enc, uint32_t ncoarse);
// The coarse level as an index file holds it (src/ixfuzz.cpp:46-143): fine
// and coarse positions from the file.
TwoLevel* two_level_from_file(const Index& ix, int enc, const std::vector<int64_t>& fine_pos,
                              const std::vector<uint32_t>& cbounds, const std::vector<int64_t>& coff,
                              const std::vector<std::vector<word_t>>& cbits_serial);
// fuzz::estimateCost / zona::estimateCost (src/ixfuzz.cpp:767-870,
// src/ixzona.cpp:580-711)
double two_level_cost(const Index& ix, const TwoLevel& tl, int type, double lo, int lop, double hi, int rop);
// fuzz::evaluate / zona::evaluate (src/ixfuzz.cpp:872-1095, src/ixzona.cpp:
// 713-1017): the hits; *option what it did, *loads bit c for every coarse
// bitmap it reads.
BV two_level_eval(const Index& ix, const TwoLevel& tl, double lo, int lop, double hi, int rop, int* option,
                  uint64_t* loads);

} // namespace orc
