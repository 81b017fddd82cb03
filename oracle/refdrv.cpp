// Thin C driver over the real FastBit library (built from /root/reference by
// oracle/ref.mk into oracle/_ref/).  TEST INFRASTRUCTURE ONLY: it pins the
// oracle restatement and writes the golden fixtures; nothing in the product
// links it.  Every function calls the reference's own public C++ API.
#include "ibis.h"
#include "irelic.h"
#include "ibin.h"
#include "iapi.h"
#include <cstdlib>
#include <cstring>
#include <vector>

static bool g_inited = false;
static void ensure_init() {
    if (!g_inited) {
        ibis::gVerbose = -1;
        ibis::init();
        g_inited = true;
    }
}

static int hand_out(const uint32_t* p, size_t n, uint32_t** out, size_t* nout) {
    *nout = n;
    *out = (uint32_t*)std::malloc(sizeof(uint32_t) * (n ? n : 1));
    if (!*out) return -1;
    if (n) std::memcpy(*out, p, sizeof(uint32_t) * n);
    return 0;
}

// Deep copy of a serialized bitvector into a reference bitvector.
static void load(ibis::bitvector& bv, const uint32_t* a, size_t n) {
    ibis::array_t<ibis::bitvector::word_t> arr(n);
    if (n) std::memcpy(arr.begin(), a, n * sizeof(uint32_t));
    ibis::bitvector tmp(arr);
    bv.copy(tmp);
}

extern "C" {

void ref_free(void* p) { std::free(p); }

int ref_wah_op(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
               int op, int compress, uint32_t** out, size_t* nout) {
    ensure_init();
    try {
        ibis::bitvector x, y;
        load(x, a, na);
        load(y, b, nb);
        switch (op) {
        case 0: x &= y; break;
        case 1: x |= y; break;
        case 2: x ^= y; break;
        default: x -= y; break;
        }
        if (compress) x.compress();
        ibis::array_t<ibis::bitvector::word_t> arr;
        x.write(arr);
        return hand_out(arr.begin(), arr.size(), out, nout);
    } catch (...) {
        return -2;
    }
}

int ref_wah_from_bits(const uint8_t* bits, size_t n, uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    for (size_t i = 0; i < n; ++i) x += (bits[i] ? 1 : 0);
    ibis::array_t<ibis::bitvector::word_t> arr;
    x.write(arr);
    return hand_out(arr.begin(), arr.size(), out, nout);
}

int ref_wah_cnt(const uint32_t* a, size_t na, uint32_t* cnt, uint32_t* size,
                uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    load(x, a, na);
    *size = x.size();
    *cnt = x.cnt();
    ibis::array_t<ibis::bitvector::word_t> arr;
    x.write(arr);
    return hand_out(arr.begin(), arr.size(), out, nout);
}

int ref_wah_compress(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    load(x, a, na);
    x.compress();
    ibis::array_t<ibis::bitvector::word_t> arr;
    x.write(arr);
    return hand_out(arr.begin(), arr.size(), out, nout);
}

int ref_wah_decompress(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    load(x, a, na);
    x.decompress();
    ibis::array_t<ibis::bitvector::word_t> arr;
    x.write(arr);
    return hand_out(arr.begin(), arr.size(), out, nout);
}

int ref_wah_flip(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    load(x, a, na);
    (void) x.size();
    x.flip();
    ibis::array_t<ibis::bitvector::word_t> arr;
    x.write(arr);
    return hand_out(arr.begin(), arr.size(), out, nout);
}

int ref_wah_rids(const uint32_t* a, size_t na, uint32_t** out, size_t* nout) {
    ensure_init();
    ibis::bitvector x;
    load(x, a, na);
    std::vector<uint32_t> r;
    for (ibis::bitvector::indexSet is = x.firstIndexSet(); is.nIndices() > 0; ++is) {
        const ibis::bitvector::word_t* ii = is.indices();
        if (is.isRange()) {
            for (ibis::bitvector::word_t k = ii[0]; k < ii[1]; ++k) r.push_back(k);
        } else {
            for (unsigned k = 0; k < is.nIndices(); ++k) r.push_back(ii[k]);
        }
    }
    return hand_out(r.data(), r.size(), out, nout);
}

} // extern "C"
