// C ABI of the bitvector part of libfastbit_hip.so (include/fastbit_hip.h).
//
// The host side mirrors the reference's operator dispatch
// (src/bitvector.cpp:1194-1762): sizes are equalised with adjustSize, the
// literal-only tests ca/cb, the all0s/all1s shortcuts and the decompressed-
// result threshold (|A|w + |B|w) * 31 >= B.nbits choose the routine, and the
// routine decides how the result words look after compress():
//   c0 / d1 / d2 / c1   -> canonical WAH of the result bits
//   c2                  -> the words the *_c2 loop emits, then compress()
//   shortcuts           -> one operand's words (compressed), active adjusted
// All word work runs in gfx950 kernels; the host only keeps the metadata the
// reference keeps in the object (nbits, nset, active word) plus host copies
// of the first and last code word for the O(1) tests and the tail appends.
#include "../../include/fastbit_hip.h"
#include "runtime.h"
#include "launch.h"
#include "bitvector_impl.h"
#include "index_dev.h"
#include "wah_core.h"
#include <cstring>
#include <vector>
#include <string>

using namespace fbh_rt;

namespace {

constexpr uint32_t MAXBITS = 31;
constexpr uint32_t ALLONES = 0x7FFFFFFFu;
constexpr uint32_t MAXCNT = 0x3FFFFFFFu;
constexpr uint32_t FILLBIT = 0x40000000u;
constexpr uint32_t HEADER0 = 0x80000000u;
constexpr uint32_t HEADER1 = 0xC0000000u;

inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
inline uint32_t popc(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

int fail(int code, const std::string& msg) { set_error(msg); return code; }

} // namespace

// ---------------------------------------------------------------------------
// Buffer management of a handle.
namespace fbh_impl {

int bv_reserve(fbh_bitvector* bv, uint32_t words) {
    if (words <= bv->cap && bv->d) return FBH_OK;
    uint32_t* nd = (uint32_t*)dev_alloc(sizeof(uint32_t) * (size_t)(words ? words : 1));
    if (!nd) return fail(FBH_ERR_NOMEM, "device allocation failed");
    if (bv->d && bv->nw) {
        int rc = check(hipMemcpyAsync(nd, bv->d, sizeof(uint32_t) * bv->nw,
                                      hipMemcpyDeviceToDevice, stream()), "copy");
        if (rc) { dev_free(nd); return rc; }
    }
    dev_free(bv->d);
    bv->d = nd;
    bv->cap = words ? words : 1;
    return FBH_OK;
}

void bv_release(fbh_bitvector* bv) {
    dev_free(bv->d); bv->d = nullptr; bv->cap = 0;
    dev_free(bv->ck); bv->ck = nullptr; bv->ckcap = 0; bv->ck_ok = false;
}

// Chunk index of the words (start group of every 32nd code word).
int bv_ensure_ck(fbh_bitvector* bv) {
    if (bv->ck_ok) return FBH_OK;
    const uint32_t nc = cdiv(bv->nw, fbk::CHUNK_WORDS);
    if (nc + 2 > bv->ckcap || !bv->ck) {
        dev_free(bv->ck);
        bv->ck = (uint32_t*)dev_alloc(sizeof(uint32_t) * (nc + 2));
        if (!bv->ck) return fail(FBH_ERR_NOMEM, "device allocation failed");
        bv->ckcap = nc + 2;
    }
    uint32_t* tmp = (uint32_t*)scratch(1).get(sizeof(uint32_t) * (nc + fbk::scan_tmp_words(nc) + 4));
    if (!tmp) return fail(FBH_ERR_NOMEM, "scratch allocation failed");
    int rc = check(fbk::launch_chunk_index(bv->d, bv->nw, bv->ck, tmp, stream()), "chunk index");
    if (rc) return rc;
    bv->ck_ok = true;
    return FBH_OK;
}

// do_cnt (src/bitvector.cpp:583-608): counts groups and ones on the device;
// an all-zero vector of more than one word collapses to one 0-fill.
int bv_do_cnt(fbh_bitvector* bv, uint32_t* nb_out) {
    uint32_t* ctl = (uint32_t*)scratch(2).get(4ull * fbk::cnt_scratch_words());
    if (!ctl) return fail(FBH_ERR_NOMEM, "scratch allocation failed");
    int rc = check(fbk::launch_cnt(bv->d, bv->nw, ctl, stream()), "cnt");
    if (rc) return rc;
    uint32_t h[2];
    rc = check(hipMemcpyAsync(h, ctl, sizeof(h), hipMemcpyDeviceToHost, stream()), "cnt copy");
    if (rc) return rc;
    rc = check(hipStreamSynchronize(stream()), "cnt sync");
    if (rc) return rc;
    const uint32_t ones = h[0], groups = h[1];
    const uint32_t nb = groups * MAXBITS;
    if (ones == 0 && bv->nw > 1) {
        const uint32_t w = HEADER0 + groups;
        rc = check(hipMemcpyAsync(bv->d, &w, 4, hipMemcpyHostToDevice, stream()), "collapse");
        if (rc) return rc;
        rc = check(hipStreamSynchronize(stream()), "collapse sync");
        if (rc) return rc;
        bv->nw = 1; bv->w0 = bv->wl = w; bv->ck_ok = false;
    }
    bv->nset = ones;
    *nb_out = nb;
    return FBH_OK;
}

int bv_size(fbh_bitvector* bv, uint32_t* out) {
    if (bv->nbits == 0 && bv->nw > 0) {
        uint32_t nb;
        int rc = bv_do_cnt(bv, &nb);
        if (rc) return rc;
        bv->nbits = nb;
    }
    *out = bv->nbits + bv->anb;
    return FBH_OK;
}

bool all0s(const fbh_bitvector* bv) {
    if (bv->nw == 0) return true;
    if (bv->nw == 1) return bv->w0 == 0 || (bv->w0 >= HEADER0 && bv->w0 < HEADER1);
    return false;
}

bool all1s(const fbh_bitvector* bv) {
    if (bv->nw == 1) return bv->w0 == ALLONES || bv->w0 > HEADER1;
    return false;
}

// ---- tail appends (src/bitvector.h:699-831) on the host copy of the last
// word; new words are collected and uploaded together.
struct Tail {
    fbh_bitvector* bv;
    std::vector<uint32_t> words;  // words[0] replaces the last word if keep
    bool has_last;
    void begin() {
        words.clear();
        has_last = bv->nw > 0;
        if (has_last) words.push_back(bv->wl);
    }
    uint32_t& back() { return words.back(); }
    bool empty() const { return words.empty(); }
    void append_active() {
        const uint32_t v = bv->aval;
        if (empty()) words.push_back(v);
        else if (v == 0) {
            if (back() == 0) back() = HEADER0 + 2;
            else if (back() >= HEADER0 && back() < HEADER1) ++back();
            else words.push_back(v);
        } else if (v == ALLONES) {
            if (back() == ALLONES) back() = HEADER1 | 2;
            else if (back() >= HEADER1) ++back();
            else words.push_back(v);
        } else {
            words.push_back(v);
        }
        bv->nbits += MAXBITS;
        bv->aval = 0; bv->anb = 0;
        bv->nset = 0;
    }
    void append_counter(int val, uint32_t cnt) {
        const uint32_t head = 2 + val;
        const uint32_t w = (head << 30) + cnt;
        bv->nbits += cnt * MAXBITS;
        if (empty()) words.push_back(w);
        else if ((back() >> 30) == head) back() += cnt;
        else if (back() == ALLONES && head == 3) back() = w + 1;
        else if (back() == 0 && head == 2) back() = w + 1;
        else words.push_back(w);
    }
    void appendFill(int val, uint32_t n) {
        if (n == 0) return;
        if (bv->anb > 0) {
            uint32_t tmp = MAXBITS - bv->anb;
            if (tmp > n) tmp = n;
            bv->anb += tmp;
            bv->aval <<= tmp;
            n -= tmp;
            if (val != 0) bv->aval |= (1U << tmp) - 1;
            if (bv->anb >= MAXBITS) append_active();
        }
        if (n >= MAXBITS) {
            const uint32_t cnt = n / MAXBITS;
            if (cnt > 1) append_counter(val, cnt);
            else { bv->aval = val ? ALLONES : 0u; append_active(); }
            n -= cnt * MAXBITS;
        }
        if (n > 0) {
            bv->anb = n;
            bv->aval = val * ((1U << n) - 1);
        }
    }
    int commit() {
        const uint32_t keep = has_last ? bv->nw - 1 : 0;
        const uint32_t nn = keep + (uint32_t)words.size();
        int rc = bv_reserve(bv, nn + 1);
        if (rc) return rc;
        if (!words.empty()) {
            rc = check(hipMemcpyAsync(bv->d + keep, words.data(), 4 * words.size(),
                                      hipMemcpyHostToDevice, stream()), "tail upload");
            if (rc) return rc;
            rc = check(hipStreamSynchronize(stream()), "tail sync");
            if (rc) return rc;
        }
        bv->nw = nn;
        if (nn) { bv->wl = words.back(); if (keep == 0) bv->w0 = words.front(); }
        bv->ck_ok = false;
        return FBH_OK;
    }
};

int bv_adjust(fbh_bitvector* bv, uint32_t nv, uint32_t nt) {
    if (bv->nbits == 0 || bv->nbits < bv->nw * MAXBITS) {
        uint32_t nb;
        int rc = bv_do_cnt(bv, &nb);
        if (rc) return rc;
        bv->nbits = nb;
    }
    const uint32_t sz = bv->nbits + bv->anb;
    if (sz == nt) return FBH_OK;
    if (nt < sz) return fail(FBH_ERR_SIZE, "adjustSize: truncation is not supported");
    Tail t{bv, {}, false};
    t.begin();
    if (nv > nt) nv = nt;
    if (nv > sz) {
        t.appendFill(1, nv - sz);
        if (nt > nv) t.appendFill(0, nt - nv);
    } else {
        t.appendFill(0, nt - sz);
    }
    return t.commit();
}

// Copies `src`'s words (and chunk index) into `dst` (metadata not touched).
int bv_copy_words(fbh_bitvector* dst, const fbh_bitvector* src) {
    int rc = bv_reserve(dst, src->nw + 1);
    if (rc) return rc;
    if (src->nw) {
        rc = check(hipMemcpyAsync(dst->d, src->d, 4ull * src->nw, hipMemcpyDeviceToDevice,
                                  stream()), "copy words");
        if (rc) return rc;
    }
    dst->nw = src->nw; dst->w0 = src->w0; dst->wl = src->wl;
    dst->ck_ok = false;
    if (src->ck_ok) {
        const uint32_t nc = cdiv(src->nw, fbk::CHUNK_WORDS) + 1;
        if (dst->ckcap < nc + 1 || !dst->ck) {
            dev_free(dst->ck);
            dst->ck = (uint32_t*)dev_alloc(4ull * (nc + 1));
            if (!dst->ck) return fail(FBH_ERR_NOMEM, "device allocation failed");
            dst->ckcap = nc + 1;
        }
        rc = check(hipMemcpyAsync(dst->ck, src->ck, 4ull * nc, hipMemcpyDeviceToDevice, stream()),
                   "copy ck");
        if (rc) return rc;
        dst->ck_ok = true;
    }
    return FBH_OK;
}

// Output of a word kernel: fresh buffers sized for the worst case.
struct OutBuf {
    uint32_t* d = nullptr; uint32_t cap = 0;
    uint32_t* ck = nullptr; uint32_t ckcap = 0;
    int alloc(uint32_t words) {
        cap = words + 2;
        ckcap = cdiv(cap, fbk::CHUNK_WORDS) + 2;
        d = (uint32_t*)dev_alloc(4ull * cap);
        ck = (uint32_t*)dev_alloc(4ull * ckcap);
        if (!d || !ck) { dev_free(d); dev_free(ck); d = ck = nullptr;
            return fail(FBH_ERR_NOMEM, "device allocation failed"); }
        return FBH_OK;
    }
};

// Reads the control block after a word kernel and installs the output.
int finish_words(fbh_bitvector* bv, OutBuf& ob, uint32_t* ctl, uint32_t* ones_out) {
    int rc = check(fbk::launch_edge_words(ob.d, ctl, stream()), "edge words");
    if (rc) return rc;
    uint32_t h[8];
    rc = check(hipMemcpyAsync(h, ctl, sizeof(h), hipMemcpyDeviceToHost, stream()), "ctl copy");
    if (rc) return rc;
    rc = check(hipStreamSynchronize(stream()), "op sync");
    if (rc) return rc;
    dev_free(bv->d); dev_free(bv->ck);
    bv->d = ob.d; bv->cap = ob.cap; bv->ck = ob.ck; bv->ckcap = ob.ckcap;
    bv->nw = h[1]; bv->w0 = h[4]; bv->wl = h[5];
    bv->ck_ok = true;
    if (ones_out) *ones_out = h[2];
    return FBH_OK;
}

// compress() (src/bitvector.cpp:378-469) of the handle's words.
int bv_compress(fbh_bitvector* bv) {
    if (bv->nw < 2) return FBH_OK;
    OutBuf ob;
    int rc = ob.alloc(bv->nw);
    if (rc) return rc;
    const uint32_t nt = fbk::tok_tiles(bv->nw);
    uint8_t* sc = (uint8_t*)scratch(0).get(64 + 16ull * nt + 64);
    if (!sc) return fail(FBH_ERR_NOMEM, "scratch allocation failed");
    uint32_t* ctl = (uint32_t*)sc;
    uint64_t* status = (uint64_t*)(sc + 64);
    rc = check(fbk::launch_tok(bv->d, nullptr, bv->nw, 0, 0, 0, ob.d, ob.ck, status, ctl,
                               stream()), "compress");
    if (rc) return rc;
    rc = finish_words(bv, ob, ctl, nullptr);
    bv->ck_ok = false;      // the compress kernel does not track group positions
    return rc;
}

// Canonical WAH (or the exact c2 words) of a OP b into `res` (a, b have
// equal group counts; `groups` > 0).
int launch_word_op(fbh_bitvector* a, fbh_bitvector* b, int op, bool c2, uint32_t groups,
                   OutBuf& ob, uint32_t** ctl_out, int c2_mode = 0) {
    if (groups == 0 && (a->nw > 1 || b->nw > 1))
        return fail(FBH_ERR_INTERNAL, "operand sizes not counted before the word op");
    const bool la = a->nw == groups;
    const bool lb = b->nw == groups;
    int rc;
    if (!la || c2) { rc = bv_ensure_ck(a); if (rc) return rc; }
    if (!lb || c2) { rc = bv_ensure_ck(b); if (rc) return rc; }
    // dense c2 operands (code words at least 1/16 of the groups): the
    // canonical group-domain kernel, whose cost follows the groups rather
    // than the words, then the c2 routines' count-1 fills written in
    const bool c2dense = c2 && c2_mode == 0 && 16ull * ((uint64_t)a->nw + b->nw) >= groups;
    if (c2 && !c2dense) {
        const uint32_t nt = fbk::op_tiles(a->nw, b->nw);
        uint8_t* sc = (uint8_t*)scratch(0).get(64 + 8ull * fbk::op_status_words(nt) + 64);
        if (!sc) return fail(FBH_ERR_NOMEM, "scratch allocation failed");
        *ctl_out = (uint32_t*)sc;
        return check(fbk::launch_op(a->d, a->ck, a->nw, b->d, b->ck, b->nw, groups, op, c2_mode,
                                    ob.d, ob.ck, (uint64_t*)(sc + 64), *ctl_out, stream()), "op");
    }
    const uint32_t nt = fbk::canon_tiles(groups);
    // [ctl 64 B][status 8 B per entry][tile starts 2 x (nt + 1) x 8 B]
    const size_t ts_off = (64 + 8ull * fbk::canon_status(groups) + 15) & ~size_t(15);
    uint8_t* sc = (uint8_t*)scratch(0).get(ts_off + 16ull * (nt + 1) + 64);
    if (!sc) return fail(FBH_ERR_NOMEM, "scratch allocation failed");
    *ctl_out = (uint32_t*)sc;
    int rc2 = check(fbk::launch_canon(a->d, a->ck, a->nw, la ? 1 : 0, b->d, b->ck, b->nw, lb ? 1 : 0,
                                      groups, op, ob.d, ob.ck, (uint64_t*)(sc + 64), *ctl_out,
                                      (uint2*)(sc + ts_off), stream()), "canonical op");
    if (!rc2 && c2dense)
        rc2 = check(fbk::launch_c2_fix(a->d, a->ck, a->nw, b->d, b->ck, b->nw, ob.d, ob.ck, *ctl_out,
                                       a->nw + b->nw, op, stream()), "c2 representations");
    return rc2;
}

int run_word_op(const fbh_bitvector* a, const fbh_bitvector* b, int op, bool c2,
                fbh_bitvector* res, uint32_t groups, int c2_mode = 0) {
    OutBuf ob;
    // c2 emits at most one word per consumed code word; a canonical result
    // has at most one word per group
    const uint32_t inw = a->nw + b->nw;
    int rc = ob.alloc(c2 ? inw : (groups > inw ? groups : inw));
    if (rc) return rc;
    uint32_t* ctl = nullptr;
    rc = launch_word_op(const_cast<fbh_bitvector*>(a), const_cast<fbh_bitvector*>(b), op, c2,
                        groups, ob, &ctl, c2_mode);
    if (rc) { dev_free(ob.d); dev_free(ob.ck); return rc; }
    return finish_words(res, ob, ctl, nullptr);
}

uint32_t active_op(uint32_t x, uint32_t y, int op) {
    switch (op) {
    case FBH_AND: return x & y;
    case FBH_OR: return x | y;
    case FBH_XOR: return x ^ y;
    default: return x & ~y;
    }
}

// a OP= b; a.compress()  -- the dispatch of src/bitvector.cpp:1194-1762.
// When `plan_only` is set the routine is decided (and the shortcut / size
// work done) but the word kernel is not run; *need_kernel tells the caller.
int bv_op(fbh_bitvector* a, const fbh_bitvector* b_in, int op, int* path_out) {
    if (op < 0 || op > 3) return fail(FBH_ERR_ARG, "unknown operator");
    uint32_t sa, sb;
    fbh_bitvector* bm = const_cast<fbh_bitvector*>(b_in);
    int rc = bv_size(a, &sa);
    if (rc) return rc;
    rc = bv_size(bm, &sb);
    if (rc) return rc;
    fbh_bitvector tmpb;
    const fbh_bitvector* b = b_in;
    if (sa > sb) {
        // the reference copies rhs, extends it and recurses
        tmpb.nbits = bm->nbits; tmpb.nset = bm->nset; tmpb.aval = bm->aval; tmpb.anb = bm->anb;
        rc = bv_copy_words(&tmpb, bm);
        if (rc) { bv_release(&tmpb); return rc; }
        rc = bv_adjust(&tmpb, 0, sa);
        if (rc) { bv_release(&tmpb); return rc; }
        b = &tmpb;
    } else if (sa < sb) {
        rc = bv_adjust(a, 0, sb);
        if (rc) return rc;
    }
    const bool ca = (a->nw * MAXBITS == a->nbits && a->nbits > 0);
    const bool cb = (b->nw * MAXBITS == b->nbits && b->nbits > 0);
    const uint32_t dsum = (a->nw + b->nw) * MAXBITS;
    const uint32_t groups = a->nbits / MAXBITS;
    int path = 0;
    enum { K_NONE, K_CANON, K_C2, K_COPYB, K_FLIPB } kind = K_NONE;

    auto active_and = [&]() {
        if (a->anb == b->anb) a->aval &= b->aval;
        else if (a->anb > b->anb) a->aval &= (b->aval << (a->anb - b->anb));
        else { a->aval = (a->aval << (b->anb - a->anb)) & b->aval; a->anb = b->anb; }
    };

    switch (op) {
    case FBH_AND:
        if (ca || cb) { kind = K_CANON; path = ca ? (cb ? FBH_PATH_C0 : FBH_PATH_D1) : FBH_PATH_D1SWAP; }
        else if (all0s(a) || all1s(b)) { active_and(); path = FBH_PATH_SHORTCUT; }
        else if (all0s(b) || all1s(a)) { kind = K_COPYB; a->nset = b->nset; active_and(); path = FBH_PATH_SHORTCUT; }
        else if (dsum >= b->nbits) { kind = K_CANON; path = FBH_PATH_D2; }
        else { kind = K_C2; path = FBH_PATH_C2; }
        break;
    case FBH_OR:
        if (ca || cb) { kind = K_CANON; path = ca ? (cb ? FBH_PATH_C0 : FBH_PATH_D1) : FBH_PATH_D1SWAP; }
        else if (all1s(a) || all0s(b)) { a->aval |= b->aval; path = FBH_PATH_SHORTCUT; }
        else if (all0s(a) || all1s(b)) { kind = K_COPYB; a->nset = b->nset; a->aval |= b->aval; path = FBH_PATH_SHORTCUT; }
        else if (dsum >= b->nbits) { kind = K_CANON; path = FBH_PATH_D2; }
        else { kind = K_C2; path = FBH_PATH_C2; }
        break;
    case FBH_XOR:
        if (ca) { kind = K_CANON; path = cb ? FBH_PATH_C0 : FBH_PATH_D1; }
        else if (cb) { kind = K_CANON; path = FBH_PATH_C1; }
        else if (dsum >= b->nbits) { kind = K_CANON; path = FBH_PATH_D2; }
        else { kind = K_C2; path = FBH_PATH_C2; }
        break;
    default:
        if (ca) { kind = K_CANON; path = cb ? FBH_PATH_C0 : FBH_PATH_D1; }
        else if (cb) { kind = K_CANON; path = FBH_PATH_C1; }
        else if (all0s(a) || all0s(b)) { a->aval &= ~b->aval; path = FBH_PATH_SHORTCUT; }
        else if (all1s(b)) {
            // zero out m_vec: one 0-fill of rhs's fill count
            a->nset = 0; a->nbits = 0; a->nw = 0; a->ck_ok = false;
            a->aval &= ~b->aval;
            Tail t{a, {}, false};
            t.begin();
            t.append_counter(0, b->w0 & MAXCNT);
            rc = t.commit();
            if (rc) { bv_release(&tmpb); return rc; }
            path = FBH_PATH_SHORTCUT;
        } else if (all1s(a)) { kind = K_FLIPB; path = FBH_PATH_SHORTCUT; }
        else if (dsum >= b->nbits) { kind = K_CANON; path = FBH_PATH_D2; }
        else { kind = K_C2; path = FBH_PATH_C2; }
        break;
    }

    if (kind == K_CANON || kind == K_C2) {
        if (groups == 0) {
            // no whole words: only the active words combine
            a->aval = active_op(a->aval, b->aval, op);
        } else {
            rc = run_word_op(a, b, op, kind == K_C2, a, groups);
            if (rc) { bv_release(&tmpb); return rc; }
            a->aval = active_op(a->aval, b->aval, op);
            a->nset = 0;
            a->nbits = groups * MAXBITS;
        }
    } else if (kind == K_COPYB) {
        rc = bv_copy_words(a, b);
        if (rc) { bv_release(&tmpb); return rc; }
        rc = bv_compress(a);
        if (rc) { bv_release(&tmpb); return rc; }
    } else if (kind == K_FLIPB) {
        const uint32_t keep = a->aval;
        rc = bv_copy_words(a, b);
        if (rc) { bv_release(&tmpb); return rc; }
        a->nbits = b->nbits; a->nset = b->nset; a->aval = b->aval; a->anb = b->anb;
        rc = check(fbk::launch_flip(a->d, a->nw, stream()), "flip");
        if (rc) { bv_release(&tmpb); return rc; }
        auto fl = [](uint32_t w) { return w > ALLONES ? (w ^ FILLBIT) : (w ^ ALLONES); };
        a->w0 = fl(a->w0); a->wl = fl(a->wl);
        if (a->nset > 0) a->nset = a->nbits - a->nset;
        if (a->anb > 0) a->aval ^= ((1U << a->anb) - 1);
        a->aval &= keep;
        rc = bv_compress(a);
        if (rc) { bv_release(&tmpb); return rc; }
    } else {
        rc = bv_compress(a);
        if (rc) { bv_release(&tmpb); return rc; }
    }
    bv_release(&tmpb);
    if (path_out) *path_out = path;
    return FBH_OK;
}


// ---------------------------------------------------------------------------
// Raw results: the words a routine writes, not compressed (the out-of-place
// operators, and the in-place operator they fall back on for operands of
// different sizes).

// Installs d[0, nw) (capacity cap) as the handle's words.
int install_words(fbh_bitvector* bv, uint32_t* d, uint32_t cap, uint32_t nw) {
    uint32_t e[2] = {0, 0};
    int rc = FBH_OK;
    if (nw) {
        rc = check(hipMemcpyAsync(&e[0], d, 4, hipMemcpyDeviceToHost, stream()), "edge copy");
        if (!rc) rc = check(hipMemcpyAsync(&e[1], d + nw - 1, 4, hipMemcpyDeviceToHost, stream()), "edge copy");
        if (!rc) rc = check(hipStreamSynchronize(stream()), "edge sync");
        if (rc) { dev_free(d); return rc; }
    }
    dev_free(bv->d); dev_free(bv->ck);
    bv->d = d; bv->cap = cap; bv->nw = nw; bv->ck = nullptr; bv->ckcap = 0; bv->ck_ok = false;
    bv->w0 = e[0]; bv->wl = e[1];
    return FBH_OK;
}

// One literal word per group of a handle (groups words, then `extra` zero
// words), in a new device buffer.
int literal_groups(fbh_bitvector* bv, uint32_t groups, uint32_t extra, uint32_t** out) {
    uint32_t* d = (uint32_t*)dev_alloc(4ull * (groups + extra + 1));
    if (!d) return fail(FBH_ERR_NOMEM, "device allocation failed");
    int rc = FBH_OK;
    if (extra) rc = check(fbk::launch_zero_u32(d + groups, extra, stream()), "zero");
    if (!rc && groups) {
        if (bv->nw == groups) {
            rc = check(hipMemcpyAsync(d, bv->d, 4ull * groups, hipMemcpyDeviceToDevice, stream()), "copy");
        } else {
            rc = bv_ensure_ck(bv);
            if (!rc) rc = check(fbk::launch_decompress(bv->d, bv->ck, bv->nw, groups, d, 0, stream()), "decompress");
        }
    }
    if (rc) { dev_free(d); return rc; }
    *out = d;
    return FBH_OK;
}

// a := a OP b as a literal array (*_c0, *_d1, *_d2).
int raw_literal(fbh_bitvector* a, fbh_bitvector* b, int op, uint32_t groups) {
    uint32_t *la = nullptr, *lb = nullptr;
    int rc = literal_groups(a, groups, 0, &la);
    if (!rc) rc = literal_groups(b, groups, 0, &lb);
    if (!rc) rc = check(fbk::launch_lit_op(la, lb, groups, op, la, stream()), "literal op");
    dev_free(lb);
    if (rc) { dev_free(la); return rc; }
    return install_words(a, la, groups + 1, groups);
}

// The words of a *_c1 loop (x compressed, y literal, both of `groups`
// groups) into res: per group value and word start, then the words.
int raw_c1(fbh_bitvector* x, fbh_bitvector* y, int rk, uint32_t groups, fbh_bitvector* res) {
    int rc = bv_ensure_ck(x);
    if (rc) return rc;
    uint32_t* V = (uint32_t*)dev_alloc(4ull * groups);
    uint32_t* H = (uint32_t*)dev_alloc(4ull * groups);
    uint32_t* O = (uint32_t*)dev_alloc(4ull * (groups + fbk::scan_tmp_words(groups)));
    if (!V || !H || !O) { dev_free(V); dev_free(H); dev_free(O); return fail(FBH_ERR_NOMEM, "device allocation failed"); }
    rc = check(fbk::launch_c1_heads(x->d, x->ck, x->nw, y->d, groups, rk, V, H, stream()), "c1 heads");
    uint32_t* tmp = (uint32_t*)dev_alloc(4ull * (fbk::scan_tmp_words(groups) + 4));
    if (!tmp && !rc) rc = fail(FBH_ERR_NOMEM, "device allocation failed");
    if (!rc) rc = check(fbk::launch_scan_excl(H, O, groups, tmp, stream()), "scan");
    uint32_t last[2] = {0, 0};
    if (!rc) rc = check(hipMemcpyAsync(&last[0], O + groups - 1, 4, hipMemcpyDeviceToHost, stream()), "copy");
    if (!rc) rc = check(hipMemcpyAsync(&last[1], H + groups - 1, 4, hipMemcpyDeviceToHost, stream()), "copy");
    if (!rc) rc = check(hipStreamSynchronize(stream()), "sync");
    const uint32_t nw = last[0] + last[1];
    uint32_t* out = nullptr;
    uint32_t* HP = nullptr;
    if (!rc) {
        out = (uint32_t*)dev_alloc(4ull * (nw + 1));
        HP = (uint32_t*)dev_alloc(4ull * (nw + 1));
        if (!out || !HP) rc = fail(FBH_ERR_NOMEM, "device allocation failed");
    }
    if (!rc) rc = check(fbk::launch_head_words(V, H, O, groups, nw, HP, out, stream()), "c1 words");
    if (!rc) rc = check(hipStreamSynchronize(stream()), "sync");
    dev_free(V); dev_free(H); dev_free(O); dev_free(tmp); dev_free(HP);
    if (rc) { dev_free(out); return rc; }
    return install_words(res, out, nw + 1, nw);
}

// The *_c1 routine `rk` with x compressed and y literal into res (x and y
// of `groups` groups), including the one-word cases of and_c1 / or_c1 /
// minus_c1 (src/bitvector.cpp:2548-2563, 2984-2999, 3751-3769), which copy
// one operand's words.
int raw_c1_routine(fbh_bitvector* x, fbh_bitvector* y, int rk, uint32_t groups, fbh_bitvector* res) {
    if (x->nw == 1 && (rk == fbk::C1_AND || rk == fbk::C1_OR || rk == fbk::C1_MINUS)) {
        const uint32_t w = x->w0;
        const bool one = w > HEADER1, zero = w > ALLONES && !one;
        fbh_bitvector* src = nullptr;
        bool flip = false;
        if (rk == fbk::C1_AND) src = one ? y : x;
        else if (rk == fbk::C1_OR) src = zero ? y : x;
        else { src = one ? y : x; flip = one; }
        if (w <= ALLONES) {
            // one literal word each (one group)
            const uint32_t yv = y->w0;
            const uint32_t v = rk == fbk::C1_AND ? (w & yv) : rk == fbk::C1_OR ? (w | yv) : (w & ~yv & ALLONES);
            uint32_t* d = (uint32_t*)dev_alloc(8);
            if (!d) return fail(FBH_ERR_NOMEM, "device allocation failed");
            int rc = check(hipMemcpyAsync(d, &v, 4, hipMemcpyHostToDevice, stream()), "upload");
            if (!rc) rc = check(hipStreamSynchronize(stream()), "sync");
            if (rc) { dev_free(d); return rc; }
            return install_words(res, d, 2, 1);
        }
        int rc = src == res ? FBH_OK : bv_copy_words(res, src);
        if (!rc && flip) {
            rc = check(fbk::launch_flip(res->d, res->nw, stream()), "flip");
            res->w0 ^= ALLONES; res->wl ^= ALLONES;      // literal words
        }
        return rc;
    }
    return raw_c1(x, y, rk, groups, res);
}

// a OP= b without compress(): the in-place dispatch (src/bitvector.cpp:
// 1194-1762) when `inplace`, else the out-of-place operator's
// (src/bitvector.cpp:1292-1407, 1488-1551, 1624-1679, 1768-1838), for
// operands of equal size; the result replaces a's words.
int bv_op_raw(fbh_bitvector* a, fbh_bitvector* b, int op, bool inplace, int* path_out) {
    const uint32_t groups = a->nbits / MAXBITS;
    const bool lit_a = a->nw * MAXBITS == a->nbits, lit_b = b->nw * MAXBITS == b->nbits;
    const bool nz = !(inplace == false && op == FBH_MINUS);        // out-of-place MINUS: no nbits > 0
    const bool ca = lit_a && (a->nbits > 0 || !nz);
    const bool cb = lit_b && (b->nbits > 0 || !nz);
    const uint32_t dsum = (a->nw + b->nw) * MAXBITS;
    const bool d2 = inplace ? dsum >= b->nbits : dsum > a->nbits;
    const uint32_t act = active_op(a->aval, b->aval, op);
    int path = 0, rc = FBH_OK;
    auto literal = [&](int p) {
        path = p;
        if (groups) rc = raw_literal(a, b, op, groups);
        a->aval = act; a->nset = 0;
    };
    auto c1 = [&](fbh_bitvector* x, fbh_bitvector* y, int rk) {
        path = FBH_PATH_C1;
        if (groups) rc = raw_c1_routine(x, y, rk, groups, a);
        a->aval = act; a->nset = 0;
    };
    auto c2 = [&]() {
        path = FBH_PATH_C2;
        if (groups) rc = run_word_op(a, b, op, true, a, groups, (int)wah::MODE_RAW);
        a->aval = act; a->nset = 0;
    };
    auto copy_b = [&]() {
        path = FBH_PATH_SHORTCUT;
        rc = bv_copy_words(a, b);
        a->nset = b->nset; a->aval = act;
    };
    const int rk_x = op == FBH_AND ? fbk::C1_AND : op == FBH_OR ? fbk::C1_OR : op == FBH_XOR ? fbk::C1_XOR
                                                                                       : fbk::C1_MINUSX;
    const int rk_y = op == FBH_MINUS ? fbk::C1_MINUS : rk_x;
    switch (op) {
    case FBH_AND:
    case FBH_OR: {
        const bool and_ = op == FBH_AND;
        if (ca && cb) literal(FBH_PATH_C0);
        else if (ca) { if (inplace) literal(FBH_PATH_D1); else c1(b, a, rk_x); }
        else if (cb) { if (inplace) literal(FBH_PATH_D1SWAP); else c1(a, b, rk_y); }
        else if (and_ ? (all0s(a) || all1s(b)) : (all1s(a) || all0s(b))) { path = FBH_PATH_SHORTCUT; a->aval = act; }
        else if (and_ ? (all0s(b) || all1s(a)) : (all0s(a) || all1s(b))) copy_b();
        else if (d2) literal(FBH_PATH_D2);
        else c2();
        break;
    }
    case FBH_XOR:
        if (ca && cb) literal(FBH_PATH_C0);
        else if (ca) { if (inplace) literal(FBH_PATH_D1); else c1(b, a, rk_x); }
        else if (cb) c1(a, b, rk_y);
        else if (d2) literal(FBH_PATH_D2);
        else c2();
        break;
    default:
        if (ca && cb) literal(FBH_PATH_C0);
        else if (ca) { if (inplace) literal(FBH_PATH_D1); else c1(b, a, rk_x); }
        else if (cb) c1(a, b, rk_y);
        else if (all0s(a) || all0s(b)) { path = FBH_PATH_SHORTCUT; a->aval = act; }
        else if (all1s(b)) {
            // one 0-fill of b's fill count (append_counter into an empty vector)
            path = FBH_PATH_SHORTCUT;
            const uint32_t w = HEADER0 | (b->w0 & MAXCNT);
            uint32_t* d = (uint32_t*)dev_alloc(8);
            if (!d) return fail(FBH_ERR_NOMEM, "device allocation failed");
            rc = check(hipMemcpyAsync(d, &w, 4, hipMemcpyHostToDevice, stream()), "upload");
            if (!rc) rc = check(hipStreamSynchronize(stream()), "sync");
            if (rc) { dev_free(d); return rc; }
            rc = install_words(a, d, 2, 1);
            a->nset = 0; a->aval = act;
        } else if (all1s(a)) {
            // ~b, active bits kept where a's are set
            path = FBH_PATH_SHORTCUT;
            const uint32_t keep = a->aval;
            rc = bv_copy_words(a, b);
            if (!rc) rc = check(fbk::launch_flip(a->d, a->nw, stream()), "flip");
            auto fl = [](uint32_t w) { return w > ALLONES ? (w ^ FILLBIT) : (w ^ ALLONES); };
            a->w0 = fl(a->w0); a->wl = fl(a->wl); a->ck_ok = false;
            a->nset = 0;
            a->aval = (b->aval ^ (b->anb ? ((1U << b->anb) - 1) : 0u)) & keep;
        }
        else if (d2) literal(FBH_PATH_D2);
        else c2();
        break;
    }
    if (rc) return rc;
    a->nbits = groups * MAXBITS;
    if (path_out) *path_out = path;
    return FBH_OK;
}

// The out-of-place operator a OP b into a new handle (src/bitvector.cpp:
// 1292-1838): operands of different sizes go through a copy of the shorter
// one extended with zeros and the in-place operator (the copy on the left,
// except for MINUS), without compress().
int bv_op_new(const fbh_bitvector* a_in, const fbh_bitvector* b_in, int op, fbh_bitvector** out, int* path) {
    if (op < 0 || op > 3) return fail(FBH_ERR_ARG, "unknown operator");
    fbh_bitvector* a = const_cast<fbh_bitvector*>(a_in);
    fbh_bitvector* b = const_cast<fbh_bitvector*>(b_in);
    uint32_t sa, sb;
    int rc = bv_size(a, &sa);
    if (!rc) rc = bv_size(b, &sb);
    if (rc) return rc;
    fbh_bitvector* res = new fbh_bitvector();
    auto copy_meta = [](fbh_bitvector* d, const fbh_bitvector* s) {
        d->nbits = s->nbits; d->nset = s->nset; d->aval = s->aval; d->anb = s->anb;
    };
    fbh_bitvector tmp;
    if (sa != sb) {
        const bool b_short = sa > sb;
        const fbh_bitvector* first = (b_short && op != FBH_MINUS) ? b : a;     // res starts as this copy
        fbh_bitvector* other = (b_short && op != FBH_MINUS) ? a : b;
        copy_meta(res, first);
        rc = bv_copy_words(res, first);
        if (!rc && first->nbits + first->anb < (sa > sb ? sa : sb)) rc = bv_adjust(res, 0, sa > sb ? sa : sb);
        if (!rc && b_short && op == FBH_MINUS) {
            // src/bitvector.cpp:1778-1782: a copy of the shorter rhs, extended
            copy_meta(&tmp, b);
            rc = bv_copy_words(&tmp, b);
            if (!rc) rc = bv_adjust(&tmp, 0, sa);
            other = &tmp;
        }
        uint32_t so = 0, sr = 0;
        if (!rc) rc = bv_size(other, &so);
        if (!rc) rc = bv_size(res, &sr);
        if (!rc) rc = bv_op_raw(res, other, op, true, path);
    } else {
        copy_meta(res, a);
        rc = bv_copy_words(res, a);
        if (!rc) rc = bv_op_raw(res, b, op, false, path);
    }
    bv_release(&tmp);
    if (rc) { bv_release(res); delete res; return rc; }
    *out = res;
    return FBH_OK;
}

// count(mask) (src/bitvector.cpp:1088-1124): the ones of a AND mask,
// computed tile by tile without writing the result.  The operands must
// have the same size.
int bv_count_mask(fbh_bitvector* a, fbh_bitvector* m, uint32_t* out) {
    uint32_t sa, sm;
    int rc = bv_size(a, &sa);
    if (!rc) rc = bv_size(m, &sm);
    if (rc) return rc;
    if (sa != sm) return fail(FBH_ERR_SIZE, "count(mask): the operands differ in size");
    const uint32_t groups = a->nbits / MAXBITS;
    const uint32_t act = popc(a->aval & m->aval);
    if (groups == 0 || all0s(a) || all0s(m)) { *out = act; return FBH_OK; }
    uint32_t c = 0;
    if (all1s(a) || all1s(m)) {
        fbh_bitvector* o = all1s(a) ? m : a;
        rc = fbh_bv_cnt(o, &c);
        if (rc) return rc;
        *out = c - popc(o->aval) + act;
        return FBH_OK;
    }
    const bool la = a->nw == groups, lm = m->nw == groups;
    if (!la && (rc = bv_ensure_ck(a))) return rc;
    if (!lm && (rc = bv_ensure_ck(m))) return rc;
    const uint32_t nb = fbk::count_and_blocks(groups);
    uint32_t* part = (uint32_t*)dev_alloc(4ull * (nb + 1));
    if (!part) return fail(FBH_ERR_NOMEM, "device allocation failed");
    rc = check(fbk::launch_count_and(a->d, a->ck, a->nw, la ? 1 : 0, m->d, m->ck, m->nw, lm ? 1 : 0, groups, part,
                                     stream()), "count");
    if (!rc) rc = check(fbk::launch_sum_parts_n(part, nb, part + nb, stream()), "count sum");
    if (!rc) rc = check(hipMemcpyAsync(&c, part + nb, 4, hipMemcpyDeviceToHost, stream()), "count copy");
    if (!rc) rc = check(hipStreamSynchronize(stream()), "count sync");
    dev_free(part);
    if (rc) return rc;
    *out = c + act;
    return FBH_OK;
}

// erase(i, j) (src/bitvector.cpp:1012-1081), with the size counted first.
// The words before the one holding bit i are kept as they are; everything
// the reference appends after them (the bits [p0, i) of that word, bits
// [j, size)) goes through append_active / append_counter and appendWord,
// which merge every token into the word before it, so it is the canonical
// compression of those bits, its first run merged into the last kept word
// when their classes match.
int bv_erase(fbh_bitvector* bv, uint32_t i, uint32_t j) {
    uint32_t size;
    int rc = bv_size(bv, &size);
    if (rc) return rc;
    if (i >= j) return FBH_OK;
    if (i >= size) return fail(FBH_ERR_ARG, "erase: the first position is past the end");
    if (j > size) j = size;
    const uint32_t groups = bv->nbits / MAXBITS;
    // the code word holding bit i: the kept prefix ends before it
    uint32_t pw = bv->nw, p0 = bv->nbits;
    if (i < bv->nbits) {
        rc = bv_ensure_ck(bv);
        if (rc) return rc;
        uint32_t ws[2] = {0, 0};
        uint32_t* dws = (uint32_t*)dev_alloc(16);
        if (!dws) return fail(FBH_ERR_NOMEM, "device allocation failed");
        rc = check(fbk::launch_word_at(bv->d, bv->ck, bv->nw, i / MAXBITS, dws, stream()), "locate");
        if (!rc) rc = check(hipMemcpyAsync(ws, dws, 8, hipMemcpyDeviceToHost, stream()), "copy");
        if (!rc) rc = check(hipStreamSynchronize(stream()), "sync");
        dev_free(dws);
        if (rc) return rc;
        pw = ws[0]; p0 = ws[1] * MAXBITS;
    }
    const uint32_t ind = i - p0;
    const uint32_t nS = ind + (size - j);
    const uint32_t whole = nS / MAXBITS, rem = nS % MAXBITS;
    // the literal groups, the active bits as one more group, a zero word
    uint32_t* L = nullptr;
    rc = literal_groups(bv, groups, 2, &L);
    if (rc) return rc;
    if (bv->anb) {
        const uint32_t act = bv->aval << (MAXBITS - bv->anb);
        rc = check(hipMemcpyAsync(L + groups, &act, 4, hipMemcpyHostToDevice, stream()), "upload");
        if (!rc) rc = check(hipStreamSynchronize(stream()), "sync");
        if (rc) { dev_free(L); return rc; }
    }
    fbh_bitvector S;                      // the appended groups, canonical
    S.d = (uint32_t*)dev_alloc(4ull * (whole + 2));
    if (!S.d) { dev_free(L); return fail(FBH_ERR_NOMEM, "device allocation failed"); }
    S.cap = whole + 2;
    rc = check(fbk::launch_shift_groups(L, p0, ind, j, nS, S.d, stream()), "shift");
    uint32_t tail = 0;
    if (!rc && rem) rc = check(hipMemcpyAsync(&tail, S.d + whole, 4, hipMemcpyDeviceToHost, stream()), "copy");
    if (!rc) rc = check(hipStreamSynchronize(stream()), "sync");
    dev_free(L);
    if (!rc && whole) {
        S.nw = whole; S.nbits = whole * MAXBITS;
        rc = check(hipMemcpy(&S.w0, S.d, 4, hipMemcpyDeviceToHost), "copy");
        if (!rc) rc = check(hipMemcpy(&S.wl, S.d + whole - 1, 4, hipMemcpyDeviceToHost), "copy");
        if (!rc) rc = bv_compress(&S);
    }
    if (rc) { bv_release(&S); return rc; }
    // stitch: kept words ++ S's words, S's first run merged into the last
    // kept word when both are all-zero or all-one groups
    auto cls = [](uint32_t w) -> int {
        if (w == 0 || (w >= HEADER0 && w < HEADER1)) return 0;
        if (w == ALLONES || w >= HEADER1) return 1;
        return -1;
    };
    auto count_of = [](uint32_t w) { return w > ALLONES ? (w & MAXCNT) : 1u; };
    uint32_t last = 0;
    if (pw) {
        rc = check(hipMemcpy(&last, bv->d + pw - 1, 4, hipMemcpyDeviceToHost), "copy");
        if (rc) { bv_release(&S); return rc; }
    }
    const bool merge = pw && S.nw && cls(last) >= 0 && cls(last) == cls(S.w0);
    const uint32_t nw = pw + S.nw - (merge ? 1u : 0u);
    uint32_t* d = (uint32_t*)dev_alloc(4ull * (nw + 1));
    if (!d) { bv_release(&S); return fail(FBH_ERR_NOMEM, "device allocation failed"); }
    if (pw) rc = check(hipMemcpyAsync(d, bv->d, 4ull * pw, hipMemcpyDeviceToDevice, stream()), "copy");
    const uint32_t skip = merge ? 1u : 0u;
    if (!rc && S.nw > skip)
        rc = check(hipMemcpyAsync(d + pw, S.d + skip, 4ull * (S.nw - skip), hipMemcpyDeviceToDevice, stream()), "copy");
    if (!rc && merge) {
        const uint32_t w = (cls(last) ? HEADER1 : HEADER0) | (count_of(last) + count_of(S.w0));
        rc = check(hipMemcpyAsync(d + pw - 1, &w, 4, hipMemcpyHostToDevice, stream()), "upload");
    }
    if (!rc) rc = check(hipStreamSynchronize(stream()), "sync");
    bv_release(&S);
    if (rc) { dev_free(d); return rc; }
    rc = install_words(bv, d, nw + 1, nw);
    if (rc) return rc;
    bv->nbits = (size - (j - i)) - rem;
    bv->nset = 0;
    bv->anb = rem;
    bv->aval = rem ? (tail >> (MAXBITS - rem)) : 0u;
    return FBH_OK;
}

} // namespace fbh_impl

using namespace fbh_impl;

// ---------------------------------------------------------------------------
extern "C" {

int fbh_init(int device) {
    FBH_API_LOCK();
    return ensure_init(device);
}
const char* fbh_last_error(void) { return last_error(); }
int fbh_sync(void) {
    FBH_API_LOCK();
    int rc = ensure_init();
    if (rc) return rc;
    return check(hipStreamSynchronize(stream()), "sync");
}

int fbh_bv_new(fbh_bitvector** out) {
    FBH_API_LOCK();
    if (!out) return fail(FBH_ERR_ARG, "null output");
    int rc = ensure_init();
    if (rc) return rc;
    *out = new fbh_bitvector();
    return FBH_OK;
}

int fbh_bv_from_serial(const uint32_t* arr, uint32_t n, fbh_bitvector** out) {
    FBH_API_LOCK();
    if (!out || (n && !arr)) return fail(FBH_ERR_ARG, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    fbh_bitvector* bv = new fbh_bitvector();
    // src/bitvector.cpp:54-105
    uint32_t nw = 0;
    if (n > 1) {
        nw = n - 1;
        const uint32_t last = arr[n - 1];
        if (last > 0 && last < MAXBITS) { bv->anb = last; bv->aval = arr[n - 2]; nw = n - 2; }
    }
    rc = bv_reserve(bv, nw + 1);
    if (rc) { delete bv; return rc; }
    if (nw) {
        rc = check(hipMemcpyAsync(bv->d, arr, 4ull * nw, hipMemcpyHostToDevice, stream()), "upload");
        if (!rc) rc = check(hipStreamSynchronize(stream()), "upload sync");
        if (rc) { bv_release(bv); delete bv; return rc; }
        bv->w0 = arr[0]; bv->wl = arr[nw - 1];
    }
    bv->nw = nw;
    *out = bv;
    return FBH_OK;
}

int fbh_bv_set(fbh_bitvector* bv, int val, uint32_t n) {
    FBH_API_LOCK();
    if (!bv) return fail(FBH_ERR_ARG, "null handle");
    // src/bitvector.cpp:256-276
    bv->nw = 0; bv->nbits = 0; bv->nset = 0; bv->aval = 0; bv->anb = 0; bv->ck_ok = false;
    Tail t{bv, {}, false};
    t.begin();
    const uint32_t k = n / MAXBITS;
    if (k > 1) t.append_counter(val ? 1 : 0, k);
    else if (k == 1) { bv->aval = val ? ALLONES : 0u; t.append_active(); }
    bv->anb = n - k * MAXBITS;
    if (val != 0) { bv->aval = (1U << bv->anb) - 1; bv->nset = k * MAXBITS; }
    return t.commit();
}

int fbh_bv_copy(const fbh_bitvector* src, fbh_bitvector** out) {
    FBH_API_LOCK();
    if (!src || !out) return fail(FBH_ERR_ARG, "null argument");
    fbh_bitvector* bv = new fbh_bitvector();
    bv->nbits = src->nbits; bv->nset = src->nset; bv->aval = src->aval; bv->anb = src->anb;
    int rc = bv_copy_words(bv, src);
    if (rc) { bv_release(bv); delete bv; return rc; }
    *out = bv;
    return FBH_OK;
}

void fbh_bv_free(fbh_bitvector* bv) {
    FBH_API_LOCK();
    if (!bv) return;
    bv_release(bv);
    delete bv;
}

int fbh_bv_write(fbh_bitvector* bv, uint32_t* arr, uint32_t cap, uint32_t* n) {
    FBH_API_LOCK();
    if (!bv || !n) return fail(FBH_ERR_ARG, "null argument");
    const uint32_t need = bv->nw + 1 + (bv->anb > 0 ? 1 : 0);
    *n = need;
    if (!arr) return FBH_OK;
    if (cap < need) return fail(FBH_ERR_SIZE, "output array too small");
    if (bv->nw) {
        int rc = check(hipMemcpyAsync(arr, bv->d, 4ull * bv->nw, hipMemcpyDeviceToHost, stream()),
                       "download");
        if (!rc) rc = check(hipStreamSynchronize(stream()), "download sync");
        if (rc) return rc;
    }
    uint32_t k = bv->nw;
    if (bv->anb > 0) arr[k++] = bv->aval;
    arr[k] = bv->anb;
    return FBH_OK;
}

int fbh_bv_op(fbh_bitvector* a, const fbh_bitvector* b, int op, int* path) {
    FBH_API_LOCK();
    if (!a || !b) return fail(FBH_ERR_ARG, "null handle");
    try { return bv_op(a, b, op, path); }
    catch (...) { return fail(FBH_ERR_INTERNAL, "exception in fbh_bv_op"); }
}

int fbh_bv_op_new(const fbh_bitvector* a, const fbh_bitvector* b, int op, fbh_bitvector** out,
                  int* path) {
    FBH_API_LOCK();
    if (!a || !b || !out) return fail(FBH_ERR_ARG, "null argument");
    try { return bv_op_new(a, b, op, out, path); }
    catch (...) { return fail(FBH_ERR_INTERNAL, "exception in fbh_bv_op_new"); }
}

int fbh_bv_count(fbh_bitvector* a, fbh_bitvector* mask, uint32_t* cnt) {
    FBH_API_LOCK();
    if (!a || !mask || !cnt) return fail(FBH_ERR_ARG, "null argument");
    return bv_count_mask(a, mask, cnt);
}

int fbh_bv_erase(fbh_bitvector* bv, uint32_t i, uint32_t j) {
    FBH_API_LOCK();
    if (!bv) return fail(FBH_ERR_ARG, "null handle");
    try { return bv_erase(bv, i, j); }
    catch (...) { return fail(FBH_ERR_INTERNAL, "exception in fbh_bv_erase"); }
}

int fbh_bv_compress(fbh_bitvector* bv) {
    FBH_API_LOCK();
    if (!bv) return fail(FBH_ERR_ARG, "null handle");
    return bv_compress(bv);
}

int fbh_bv_decompress(fbh_bitvector* bv) {
    FBH_API_LOCK();
    if (!bv) return fail(FBH_ERR_ARG, "null handle");
    // src/bitvector.cpp:474-512
    if (bv->nbits == 0 && bv->nw > 0) {
        uint32_t nb;
        int rc = bv_do_cnt(bv, &nb);
        if (rc) return rc;
        bv->nbits = nb;
    }
    if (bv->nw * MAXBITS == bv->nbits) return FBH_OK;
    const uint32_t groups = bv->nbits / MAXBITS;
    int rc = bv_ensure_ck(bv);
    if (rc) return rc;
    uint32_t* nd = (uint32_t*)dev_alloc(4ull * (groups + 1));
    if (!nd) return fail(FBH_ERR_NOMEM, "device allocation failed");
    rc = check(fbk::launch_decompress(bv->d, bv->ck, bv->nw, groups, nd, 0, stream()), "decompress");
    if (rc) { dev_free(nd); return rc; }
    rc = check(hipStreamSynchronize(stream()), "decompress sync");
    if (rc) { dev_free(nd); return rc; }
    uint32_t edge[2] = {0, 0};
    if (groups) {
        (void)hipMemcpy(&edge[0], nd, 4, hipMemcpyDeviceToHost);
        (void)hipMemcpy(&edge[1], nd + groups - 1, 4, hipMemcpyDeviceToHost);
    }
    dev_free(bv->d);
    bv->d = nd; bv->cap = groups + 1; bv->nw = groups;
    bv->w0 = edge[0]; bv->wl = edge[1];
    bv->ck_ok = false;
    return FBH_OK;
}

int fbh_bv_flip(fbh_bitvector* bv) {
    FBH_API_LOCK();
    if (!bv) return fail(FBH_ERR_ARG, "null handle");
    // src/bitvector.cpp:1128-1174
    int rc = check(fbk::launch_flip(bv->d, bv->nw, stream()), "flip");
    if (rc) return rc;
    auto fl = [](uint32_t w) { return w > ALLONES ? (w ^ FILLBIT) : (w ^ ALLONES); };
    if (bv->nw) { bv->w0 = fl(bv->w0); bv->wl = fl(bv->wl); }
    if (bv->nbits == 0 && bv->nw > 0) {
        uint32_t* ctl = (uint32_t*)scratch(2).get(4ull * fbk::cnt_scratch_words());
        if (!ctl) return fail(FBH_ERR_NOMEM, "scratch allocation failed");
        rc = check(fbk::launch_cnt(bv->d, bv->nw, ctl, stream()), "flip count");
        if (rc) return rc;
        uint32_t h[2];
        rc = check(hipMemcpyAsync(h, ctl, 8, hipMemcpyDeviceToHost, stream()), "flip copy");
        if (!rc) rc = check(hipStreamSynchronize(stream()), "flip sync");
        if (rc) return rc;
        bv->nbits = h[1] * MAXBITS;
    }
    if (bv->nset > 0) bv->nset = bv->nbits - bv->nset;
    if (bv->anb > 0) bv->aval ^= ((1U << bv->anb) - 1);
    return FBH_OK;
}

int fbh_bv_cnt(fbh_bitvector* bv, uint32_t* cnt) {
    FBH_API_LOCK();
    if (!bv || !cnt) return fail(FBH_ERR_ARG, "null argument");
    // src/bitvector.h:652-656
    if (bv->nset == 0 && bv->nw > 0) {
        uint32_t nb;
        int rc = bv_do_cnt(bv, &nb);
        if (rc) return rc;
        bv->nbits = nb;
    }
    *cnt = bv->nset + popc(bv->aval);
    return FBH_OK;
}

int fbh_bv_size(fbh_bitvector* bv, uint32_t* size) {
    FBH_API_LOCK();
    if (!bv || !size) return fail(FBH_ERR_ARG, "null argument");
    return bv_size(bv, size);
}

int fbh_bv_adjust_size(fbh_bitvector* bv, uint32_t nv, uint32_t nt) {
    FBH_API_LOCK();
    if (!bv) return fail(FBH_ERR_ARG, "null handle");
    return bv_adjust(bv, nv, nt);
}

int fbh_bv_nwords(const fbh_bitvector* bv, uint32_t* n) {
    FBH_API_LOCK();
    if (!bv || !n) return fail(FBH_ERR_ARG, "null argument");
    *n = bv->nw;
    return FBH_OK;
}

int fbh_bv_rids(fbh_bitvector* bv, uint32_t* rids, uint32_t cap, uint32_t* n) {
    FBH_API_LOCK();
    if (!bv || !n) return fail(FBH_ERR_ARG, "null argument");
    uint32_t c;
    int rc = fbh_bv_cnt(bv, &c);
    if (rc) return rc;
    *n = c;
    if (!rids) return FBH_OK;
    const uint32_t groups = bv->nbits / MAXBITS;
    const uint32_t nword_rids = c - popc(bv->aval);
    if (groups && nword_rids) {
        rc = bv_ensure_ck(bv);
        if (rc) return rc;
        const bool literal = bv->nw == groups;
        uint32_t* lit = bv->d;
        uint32_t* litbuf = nullptr;
        if (!literal) {
            litbuf = (uint32_t*)dev_alloc(4ull * groups);
            if (!litbuf) return fail(FBH_ERR_NOMEM, "device allocation failed");
            rc = check(fbk::launch_decompress(bv->d, bv->ck, bv->nw, groups, litbuf, 0, stream()),
                       "rids decompress");
            if (rc) { dev_free(litbuf); return rc; }
            lit = litbuf;
        }
        uint32_t* out = (uint32_t*)dev_alloc(4ull * nword_rids);
        uint32_t* tmp = (uint32_t*)scratch(1).get(4ull * (2 * groups + fbk::scan_tmp_words(groups) + 4));
        if (!out || !tmp) { dev_free(litbuf); dev_free(out); return fail(FBH_ERR_NOMEM, "alloc"); }
        rc = check(fbk::launch_rids_lit(lit, groups, out, nword_rids, tmp, stream()), "rids");
        const uint32_t ncopy = nword_rids < cap ? nword_rids : cap;
        if (!rc) rc = check(hipMemcpyAsync(rids, out, 4ull * ncopy, hipMemcpyDeviceToHost, stream()),
                            "rids copy");
        if (!rc) rc = check(hipStreamSynchronize(stream()), "rids sync");
        dev_free(litbuf); dev_free(out);
        if (rc) return rc;
    }
    // the active word's bits follow (first bit = bit anb-1)
    uint32_t k = nword_rids;
    const uint32_t pos = bv->nbits;
    for (uint32_t i = 0; i < bv->anb; ++i)
        if ((bv->aval >> (bv->anb - 1 - i)) & 1U) { if (k < cap) rids[k] = pos + i; ++k; }
    return FBH_OK;
}

int fbh_bench_op(const fbh_bitvector* a, const fbh_bitvector* b, int op, int reps, float* ms,
                 uint32_t* out_words) {
    FBH_API_LOCK();
    if (!a || !b || reps <= 0) return fail(FBH_ERR_ARG, "bad argument");
    // decide the routine on a throw-away copy (metadata + shortcut work)
    fbh_bitvector* ta = nullptr;
    int rc = fbh_bv_copy(a, &ta);
    if (rc) return rc;
    int path = 0;
    rc = fbh_bv_op(ta, b, op, &path);
    const uint32_t words = ta ? ta->nw : 0;
    fbh_bv_free(ta);
    if (rc) return rc;
    if (out_words) *out_words = words;
    if (path != FBH_PATH_C0 && path != FBH_PATH_D1 && path != FBH_PATH_D1SWAP &&
        path != FBH_PATH_C1 && path != FBH_PATH_D2 && path != FBH_PATH_C2) {
        if (ms) *ms = 0.f;
        return FBH_OK;
    }
    fbh_bitvector* ma = const_cast<fbh_bitvector*>(a);
    fbh_bitvector* mb = const_cast<fbh_bitvector*>(b);
    uint32_t sa = 0, sb = 0;     // nbits is computed lazily (do_cnt)
    if ((rc = bv_size(ma, &sa)) || (rc = bv_size(mb, &sb))) return rc;
    const uint32_t groups = ma->nbits / MAXBITS;
    const bool c2 = path == FBH_PATH_C2;
    OutBuf ob;
    rc = ob.alloc(ma->nw + mb->nw);
    if (rc) return rc;
    uint32_t* ctl = nullptr;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
    auto one = [&]() { return launch_word_op(ma, mb, op, c2, groups, ob, &ctl); };
    rc = one();
    if (!rc) rc = check(hipEventRecord(e0, stream()), "event");
    for (int i = 0; i < reps && !rc; ++i) rc = one();
    if (!rc) rc = check(hipEventRecord(e1, stream()), "event");
    if (!rc) rc = check(hipEventSynchronize(e1), "event sync");
    float t = 0.f;
    if (!rc) (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    dev_free(ob.d); dev_free(ob.ck);
    if (ms) *ms = t / reps;
    return rc;
}

// Like fbh_bench_op, cycling over npairs operand pairs of the same shape so
// that the working set exceeds the last-level cache and every op streams
// its operands from HBM.
int fbh_bench_cnt_rot(fbh_bitvector* const* a, int n, int reps, float* ms, uint32_t* ones) {
    FBH_API_LOCK();
    if (!a || n <= 0 || reps <= 0) return fail(FBH_ERR_ARG, "bad argument");
    uint32_t* ctl = (uint32_t*)scratch(2).get(4ull * fbk::cnt_scratch_words());
    if (!ctl) return fail(FBH_ERR_NOMEM, "scratch allocation failed");
    int rc = FBH_OK;
    for (int i = 0; i < n && !rc; ++i)
        rc = check(fbk::launch_cnt(a[i]->d, a[i]->nw, ctl, stream()), "cnt");
    uint32_t h[2] = {0, 0};
    if (!rc) rc = check(hipMemcpyAsync(h, ctl, sizeof(h), hipMemcpyDeviceToHost, stream()), "cnt copy");
    if (!rc) rc = check(hipStreamSynchronize(stream()), "cnt sync");
    if (ones) *ones = h[0];
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
    if (!rc) rc = check(hipEventRecord(e0, stream()), "event");
    for (int r = 0; r < reps && !rc; ++r)
        rc = check(fbk::launch_cnt(a[r % n]->d, a[r % n]->nw, ctl, stream()), "cnt");
    if (!rc) rc = check(hipEventRecord(e1, stream()), "event");
    if (!rc) rc = check(hipEventSynchronize(e1), "event sync");
    float t = 0.f;
    if (!rc) (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    if (ms) *ms = t / reps;
    return rc;
}

int fbh_bench_op_rot(fbh_bitvector* const* a, fbh_bitvector* const* b, int npairs, int op, int reps,
                     float* ms, uint32_t* out_words) {
    FBH_API_LOCK();
    if (!a || !b || npairs <= 0 || reps <= 0) return fail(FBH_ERR_ARG, "bad argument");
    std::vector<int> paths(npairs);
    for (int i = 0; i < npairs; ++i) {
        fbh_bitvector* ta = nullptr;
        int rc = fbh_bv_copy(a[i], &ta);
        if (rc) return rc;
        rc = fbh_bv_op(ta, b[i], op, &paths[i]);
        if (i == 0 && out_words) *out_words = ta->nw;
        fbh_bv_free(ta);
        if (rc) return rc;
        if (paths[i] == FBH_PATH_SHORTCUT) { if (ms) *ms = 0.f; return FBH_OK; }
    }
    for (int i = 0; i < npairs; ++i) {
        uint32_t sa = 0, sb = 0;     // nbits is computed lazily (do_cnt)
        int rc = bv_size(a[i], &sa);
        if (!rc) rc = bv_size(b[i], &sb);
        if (rc) return rc;
    }
    std::vector<OutBuf> obs(npairs);
    for (int i = 0; i < npairs; ++i) {
        const bool c2 = paths[i] == FBH_PATH_C2;
        const uint32_t groups = a[i]->nbits / MAXBITS;
        int rc = obs[i].alloc(c2 ? a[i]->nw + b[i]->nw : (groups > a[i]->nw + b[i]->nw ? groups : a[i]->nw + b[i]->nw));
        if (rc) return rc;
    }
    uint32_t* ctl = nullptr;
    auto one = [&](int i) {
        const uint32_t groups = a[i]->nbits / MAXBITS;
        return launch_word_op(a[i], b[i], op, paths[i] == FBH_PATH_C2, groups, obs[i], &ctl);
    };
    int rc = FBH_OK;
    for (int i = 0; i < npairs && !rc; ++i) rc = one(i);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
    if (!rc) rc = check(hipEventRecord(e0, stream()), "event");
    for (int r = 0; r < reps && !rc; ++r) rc = one(r % npairs);
    if (!rc) rc = check(hipEventRecord(e1, stream()), "event");
    if (!rc) rc = check(hipEventSynchronize(e1), "event sync");
    float t = 0.f;
    if (!rc) (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    for (auto& ob : obs) { dev_free(ob.d); dev_free(ob.ck); }
    if (ms) *ms = t / reps;
    return rc;
}

} // extern "C"
