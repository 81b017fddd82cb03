// C entry points of the CPU oracle, for the Python test harness (ctypes).
// TEST INFRASTRUCTURE ONLY -- the product library never links this.
//
// Bitvectors cross this boundary in the reference's serialized form
// (src/bitvector.cpp:2283-2313): the m_vec words, then active.val when
// active.nbits > 0, then active.nbits.
#include "wah.hpp"
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace orc;

static int hand_out(const std::vector<word_t>& s, uint32_t** out, size_t* nout) {
    *nout = s.size();
    *out = (uint32_t*)std::malloc(sizeof(uint32_t) * (s.size() ? s.size() : 1));
    if (!*out) return -1;
    if (!s.empty()) std::memcpy(*out, s.data(), sizeof(uint32_t) * s.size());
    return 0;
}

extern "C" {

void orc_free(void* p) { std::free(p); }

// a OP= b, then (optionally) a.compress(); the result is returned serialized.
// *path receives the routine that ran (orc::Path).
int orc_wah_op(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
               int op, int compress, uint32_t** out, size_t* nout, int* path) {
    try {
        BV x = BV::from_serial(a, na);
        BV y = BV::from_serial(b, nb);
        Path p = x.op_inplace(y, (Op)op);
        if (compress) x.compress();
        if (path) *path = (int)p;
        return hand_out(x.serial(), out, nout);
    } catch (...) {
        return -2;
    }
}

// One *_c2 routine forced (whatever the dispatch would pick), then compress().
int orc_wah_c2(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
               int op, uint32_t** out, size_t* nout) {
    try {
        BV x = BV::from_serial(a, na);
        BV y = BV::from_serial(b, nb);
        x.size(); y.size();
        BV res;
        switch (op) {
        case OP_AND: x.and_c2(y, res); break;
        case OP_OR: x.or_c2(y, res); break;
        case OP_XOR: x.xor_c2(y, res); break;
        default: x.minus_c2(y, res); break;
        }
        res.compress();
        return hand_out(res.serial(), out, nout);
    } catch (...) {
        return -2;
    }
}

// One *_c2 routine forced, its words as the loop emits them (no compress()).
int orc_wah_c2_raw(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
                   int op, uint32_t** out, size_t* nout) {
    try {
        BV x = BV::from_serial(a, na);
        BV y = BV::from_serial(b, nb);
        x.size(); y.size();
        BV res;
        switch (op) {
        case OP_AND: x.and_c2(y, res); break;
        case OP_OR: x.or_c2(y, res); break;
        case OP_XOR: x.xor_c2(y, res); break;
        default: x.minus_c2(y, res); break;
        }
        return hand_out(res.serial(), out, nout);
    } catch (...) {
        return -2;
    }
}

// Canonical WAH of a bit string given as bytes of 0/1 (operator+= per bit).
int orc_wah_from_bits(const uint8_t* bits, size_t n, uint32_t** out, size_t* nout) {
    BV x;
    for (size_t i = 0; i < n; ++i) x.push_bit(bits[i] ? 1 : 0);
    return hand_out(x.serial(), out, nout);
}

// Canonical WAH of a run-length description: runs alternate starting with
// value first_bit; lens[i] bits each (appendFill per run).
int orc_wah_from_runs(int first_bit, const uint32_t* lens, size_t nruns,
                      uint32_t** out, size_t* nout) {
    BV x;
    int b = first_bit ? 1 : 0;
    for (size_t i = 0; i < nruns; ++i) { x.appendFill(b, lens[i]); b ^= 1; }
    return hand_out(x.serial(), out, nout);
}

int orc_wah_compress(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    BV x = BV::from_serial(a, na);
    x.compress();
    return hand_out(x.serial(), out, nout);
}

int orc_wah_decompress(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    BV x = BV::from_serial(a, na);
    x.decompress();
    return hand_out(x.serial(), out, nout);
}

// cnt() and size(); the words after the call (do_cnt may collapse an
// all-zero vector) are returned serialized.
int orc_wah_cnt(const uint32_t* a, size_t na, uint32_t* cnt, uint32_t* size,
                uint32_t** out, size_t* nout) {
    BV x = BV::from_serial(a, na);
    *size = x.size();
    *cnt = x.cnt();
    return hand_out(x.serial(), out, nout);
}

int orc_wah_flip(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    BV x = BV::from_serial(a, na);
    x.size();
    x.flip();
    return hand_out(x.serial(), out, nout);
}

int orc_wah_rids(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    BV x = BV::from_serial(a, na);
    std::vector<word_t> r;
    x.rids(r);
    return hand_out(r, out, nout);
}

// Literal words (one 31-bit group per word, no fill words) of a bitvector.
int orc_wah_literal(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    BV x = BV::from_serial(a, na);
    return hand_out(x.literal_words(), out, nout);
}

} // extern "C"

extern "C" {

// The out-of-place operator a OP b (not compressed); *path receives the
// routine (orc::Path).
int orc_wah_op_new(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, int op,
                   uint32_t** out, size_t* nout, int* path) {
    try {
        BV x = BV::from_serial(a, na);
        BV y = BV::from_serial(b, nb);
        Path p = PATH_NONE;
        BV r = x.op_new(y, (Op)op, &p);
        if (path) *path = (int)p;
        return hand_out(r.serial(), out, nout);
    } catch (...) {
        return -2;
    }
}

// count(mask): the ones of a AND mask.
int orc_wah_count(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* cnt) {
    try {
        BV x = BV::from_serial(a, na);
        BV y = BV::from_serial(b, nb);
        x.size(); y.size();
        *cnt = x.count_mask(y);
        return 0;
    } catch (...) {
        return -2;
    }
}

// erase(i, j)
int orc_wah_erase(const uint32_t* a, size_t na, uint32_t i, uint32_t j, uint32_t** out, size_t* nout) {
    try {
        BV x = BV::from_serial(a, na);
        x.erase(i, j);
        return hand_out(x.serial(), out, nout);
    } catch (...) {
        return -2;
    }
}

} // extern "C"
