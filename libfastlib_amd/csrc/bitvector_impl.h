// Device bitvector handle (the object behind fbh_bitvector*) and the
// internal operations the index code shares with the bitvector ABI.
#pragma once
#include <stdint.h>

// Mirrors ibis::bitvector's members (src/bitvector.h:300-304) with m_vec in
// HBM.  nbits == 0 means "not counted yet" and nset == 0 "unknown", as in
// the reference, whose size()/cnt() count lazily.
struct fbh_bitvector {
    uint32_t* d = nullptr;   // code words (m_vec)
    uint32_t cap = 0;        // capacity of d (words)
    uint32_t nw = 0;         // m_vec.size()
    uint32_t* ck = nullptr;  // chunk index: start group of every 32nd word
    uint32_t ckcap = 0;
    bool ck_ok = false;
    uint32_t nbits = 0;      // bits in m_vec
    uint32_t nset = 0;       // ones in m_vec
    uint32_t aval = 0;       // active word value
    uint32_t anb = 0;        // active word bit count
    uint32_t w0 = 0;         // host copy of m_vec[0]
    uint32_t wl = 0;         // host copy of m_vec.back()
};

namespace fbh_impl {
int bv_reserve(fbh_bitvector* bv, uint32_t words);
void bv_release(fbh_bitvector* bv);
int bv_ensure_ck(fbh_bitvector* bv);
int bv_do_cnt(fbh_bitvector* bv, uint32_t* nb_out);
int bv_size(fbh_bitvector* bv, uint32_t* out);
int bv_compress(fbh_bitvector* bv);
int bv_op(fbh_bitvector* a, const fbh_bitvector* b, int op, int* path_out);
int bv_copy_words(fbh_bitvector* dst, const fbh_bitvector* src);
} // namespace fbh_impl
