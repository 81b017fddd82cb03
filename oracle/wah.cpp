// CPU oracle: restatement of ibis::bitvector (FastBit 2.0.3, src/bitvector.h,
// src/bitvector.cpp).  TEST INFRASTRUCTURE ONLY -- see wah.hpp.
#include "wah.hpp"
#include <algorithm>
#include <cstring>

namespace orc {

word_t op_word(word_t a, word_t b, Op op) {
    switch (op) {
    case OP_AND: return a & b;
    case OP_OR: return a | b;
    case OP_XOR: return a ^ b;
    default: return a & ~b;
    }
}

void literal_op(const std::vector<word_t>& a, const std::vector<word_t>& b,
                Op op, std::vector<word_t>& out) {
    if (a.size() != b.size()) throw std::runtime_error("literal_op: length mismatch");
    out.resize(a.size());
    for (size_t i = 0; i < a.size(); ++i) out[i] = op_word(a[i], b[i], op);
}

// src/bitvector.cpp:54-105: the last word is active.nbits; when it is
// non-zero the word before it is active.val.  nbits stays 0 (lazy count).
BV BV::from_serial(const word_t* a, size_t n) {
    BV r;
    if (n > 1) {
        r.v.assign(a, a + n);
        if (r.v.back() > 0) {
            if (r.v.back() < MAXBITS) {
                r.anb = r.v.back();
                r.v.pop_back();
                r.aval = r.v.back();
            }
        } else {
            r.aval = 0; r.anb = 0;
        }
        r.v.pop_back();
    } else {
        r.clear();
    }
    return r;
}

// src/bitvector.cpp:2283-2313
std::vector<word_t> BV::serial() const {
    std::vector<word_t> out(v);
    if (anb > 0) out.push_back(aval);
    out.push_back(anb);
    return out;
}

// src/bitvector.cpp:256-276
void BV::set(int val, word_t n) {
    clear();
    word_t k = n / MAXBITS;
    if (k > 1) {
        append_counter(val, k);
    } else if (k == 1) {
        aval = val ? ALLONES : 0;
        append_active();
    }
    anb = n - k * MAXBITS;
    if (val != 0) {
        aval = (1U << anb) - 1;
        nset = k * MAXBITS;
    }
}

// src/bitvector.h:699-737
void BV::append_active() {
    if (v.empty()) {
        v.push_back(aval);
    } else if (aval == 0) {
        if (v.back() == 0) v.back() = (HEADER0 + 2);
        else if (v.back() >= HEADER0 && v.back() < HEADER1) ++v.back();
        else v.push_back(aval);
    } else if (aval == ALLONES) {
        if (v.back() == ALLONES) v.back() = (HEADER1 | 2);
        else if (v.back() >= HEADER1) ++v.back();
        else v.push_back(aval);
    } else {
        v.push_back(aval);
    }
    nbits += MAXBITS;
    aval = 0; anb = 0;
    nset = 0;
}

// src/bitvector.h:741-760
void BV::append_counter(int val, word_t cnt) {
    word_t head = 2 + val;
    word_t w = (head << SECONDBIT) + cnt;
    nbits += cnt * MAXBITS;
    if (v.empty()) v.push_back(w);
    else if ((v.back() >> SECONDBIT) == head) v.back() += cnt;
    else if ((v.back() == ALLONES) && head == 3) v.back() = w + 1;
    else if ((v.back() == 0) && head == 2) v.back() = w + 1;
    else v.push_back(w);
}

// src/bitvector.h:763-768
void BV::push_bit(int b) {
    aval <<= 1; ++anb; aval += b;
    if (anb >= MAXBITS) append_active();
}

// src/bitvector.h:800-831
void BV::appendFill(int val, word_t n) {
    if (n == 0) return;
    if (anb > 0) {
        word_t tmp = (MAXBITS - anb);
        if (tmp > n) tmp = n;
        anb += tmp;
        aval <<= tmp;
        n -= tmp;
        if (val != 0) aval |= (1U << tmp) - 1;
        if (anb >= MAXBITS) append_active();
    }
    if (n >= MAXBITS) {
        word_t cnt = n / MAXBITS;
        if (cnt > 1) append_counter(val, cnt);
        else if (val != 0) { aval = ALLONES; append_active(); }
        else { aval = 0; append_active(); }
        n -= cnt * MAXBITS;
    }
    if (n > 0) {
        anb = n;
        aval = val * ((1U << n) - 1);
    }
}

// src/bitvector.cpp:621-704 (extension past the end and the active word;
// index builds only append in ascending row order).
void BV::setBit(word_t ind, int val) {
    if (ind >= size()) {
        word_t diff = ind - size() + 1;
        if (anb) {
            if (ind + 1 >= nbits + MAXBITS) {
                diff -= MAXBITS - anb;
                aval <<= (MAXBITS - anb);
                if (diff == 0) aval += (val != 0);
                append_active();
            } else {
                anb += diff;
                aval <<= diff;
                aval += (val != 0);
                diff = 0;
            }
        }
        if (diff) {
            word_t w = diff / MAXBITS;
            diff -= w * MAXBITS;
            if (diff) {
                if (w > 1) append_counter(0, w);
                else if (w) append_active();
                anb = diff;
                aval += (val != 0);
            } else if (val != 0) {
                if (w > 2) append_counter(0, w - 1);
                else if (w == 2) append_active();
                aval = 1;
                append_active();
            } else {
                if (w > 1) append_counter(0, w);
                else if (w) append_active();
            }
        }
        if (nset) nset += (val != 0);
        return;
    } else if (ind >= nbits) {
        word_t u = aval;
        if (val != 0) aval |= (1U << (anb - (ind - nbits) - 1));
        else aval &= ~(1U << (anb - (ind - nbits) - 1));
        if (nset && (u != aval)) nset += (val ? 1 : -1);
        return;
    }
    throw std::runtime_error("oracle setBit: only append and active-word updates are restated");
}

// src/bitvector.h:838-872
void BV::copy_runs(Run& it, word_t& nw) {
    if (it.isFill == 0) {
        aval = *(it.it);
        append_active();
        --nw;
    } else if (it.nWords > 1) {
        append_counter(it.fillBit, it.nWords);
        nw -= it.nWords;
    } else if (it.nWords == 1) {
        aval = (it.fillBit != 0 ? ALLONES : 0);
        append_active();
        --nw;
    }
    ++it.it;
    nset = 0;
    it.nWords = 0;
    nbits += MAXBITS * nw;
    while (nw > 0) {
        it.decode();
        if (nw >= it.nWords) {
            v.push_back(*(it.it));
            nw -= it.nWords;
            it.nWords = 0;
            ++it.it;
        } else {
            break;
        }
    }
    nbits -= MAXBITS * nw;
}

// src/bitvector.h:876-910
void BV::copy_runsn(Run& it, word_t& nw) {
    if (it.isFill == 0) {
        aval = ALLONES ^ *(it.it);
        append_active();
        --nw;
    } else if (it.nWords > 1) {
        append_counter(!it.fillBit, it.nWords);
        nw -= it.nWords;
    } else if (it.nWords == 1) {
        aval = (it.fillBit != 0 ? 0 : ALLONES);
        append_active();
        --nw;
    }
    ++it.it;
    nset = 0;
    it.nWords = 0;
    nbits += MAXBITS * nw;
    while (nw > 0) {
        it.decode();
        if (nw >= it.nWords) {
            v.push_back((it.isFill ? FILLBIT : ALLONES) ^ *(it.it));
            nw -= it.nWords;
            it.nWords = 0;
            ++it.it;
        } else {
            break;
        }
    }
    nbits -= MAXBITS * nw;
}

// src/bitvector.cpp:583-608, including the collapse of an all-zero vector
// of more than one word into a single 0-fill.
word_t BV::do_cnt() const {
    word_t ns = 0, nb = 0;
    for (size_t i = 0; i < v.size(); ++i) {
        word_t w = v[i];
        if (w < HEADER0) { nb += MAXBITS; ns += cnt_ones(w); }
        else {
            word_t tmp = (w & MAXCNT) * MAXBITS;
            nb += tmp;
            ns += tmp * (w >= HEADER1);
        }
    }
    if (ns == 0 && v.size() > 1) {
        v[0] = HEADER0 + (nb / MAXBITS);
        v.resize(1);
    }
    nset = ns;
    return nb;
}

// src/bitvector.h:652-656
word_t BV::cnt() const {
    if (nset == 0 && !v.empty()) nbits = do_cnt();
    return nset + cnt_ones(aval);
}

bool BV::all0s() const {
    if (v.empty()) return true;
    if (v.size() == 1) return (v[0] == 0 || (v[0] >= HEADER0 && v[0] < HEADER1));
    return false;
}

bool BV::all1s() const {
    if (v.size() == 1) return (v[0] == ALLONES || (v[0] > HEADER1));
    return false;
}

// src/bitvector.cpp:378-469
void BV::compress() {
    if (v.size() < 2) return;
    size_t last = 0;
    bool lastIsFill = v[0] > ALLONES;
    int lastFillBit = v[0] > HEADER1;
    for (size_t cur = 1; cur < v.size(); ++cur) {
        const word_t cw = v[cur];
        const bool curIsFill = cw > ALLONES;
        const int curFillBit = cw > HEADER1;
        const word_t curN = cw & MAXCNT;
        if (lastIsFill) {
            if (curIsFill) {
                if (curFillBit == lastFillBit) {
                    v[last] += curN;
                } else {
                    ++last; v[last] = cw; lastFillBit = curFillBit;
                }
            } else if ((lastFillBit == 0 && cw == 0) ||
                       (lastFillBit != 0 && cw == ALLONES)) {
                ++v[last];
            } else {
                ++last; lastIsFill = false; v[last] = cw;
            }
        } else if (curIsFill) {
            if ((curFillBit == 0 && v[last] == 0) ||
                (curFillBit != 0 && v[last] == ALLONES)) {
                v[last] = cw + 1;
                lastFillBit = curFillBit;
                lastIsFill = true;
            } else {
                ++last; lastIsFill = true; v[last] = cw; lastFillBit = curFillBit;
            }
        } else if (v[last] == cw) {
            if (cw == 0) { v[last] = HEADER0 | 2; lastIsFill = true; lastFillBit = 0; }
            else if (cw == ALLONES) { v[last] = HEADER1 | 2; lastIsFill = true; lastFillBit = 1; }
            else { ++last; v[last] = cw; }
        } else {
            ++last; v[last] = cw;
        }
    }
    v.resize(last + 1);
}

// src/bitvector.cpp:515-541
std::vector<word_t> BV::literal_words() const {
    const word_t nb = ((nbits == 0 && v.size() > 0) ? do_cnt() : nbits);
    std::vector<word_t> out;
    out.reserve(nb / MAXBITS);
    for (size_t i = 0; i < v.size(); ++i) {
        word_t w = v[i];
        if (w > ALLONES) {
            word_t cnt = w & MAXCNT;
            word_t f = (w >= HEADER1) ? ALLONES : 0;
            out.insert(out.end(), cnt, f);
        } else {
            out.push_back(w);
        }
    }
    return out;
}

// src/bitvector.cpp:474-512
void BV::decompress() {
    if (nbits == 0 && v.size() > 0) nbits = do_cnt();
    if (v.size() * MAXBITS == nbits) return;
    std::vector<word_t> tmp = literal_words();
    v.swap(tmp);
}

// src/bitvector.cpp:1128-1174
void BV::flip() {
    if (nbits > 0) {
        for (auto& w : v) w = (w > ALLONES) ? (w ^ FILLBIT) : (w ^ ALLONES);
    } else {
        nbits = 0;
        for (auto& w : v) {
            if (w > ALLONES) { w ^= FILLBIT; nbits += MAXBITS * (w & MAXCNT); }
            else { w ^= ALLONES; nbits += MAXBITS; }
        }
    }
    if (nset > 0) nset = nbits - nset;
    if (anb > 0) aval ^= ((1U << anb) - 1);
}

// src/bitvector.cpp:4086-4113 (growth only; truncation goes through erase,
// which no path restated here reaches).
void BV::adjustSize(word_t nv, word_t nt) {
    if (nbits == 0 || nbits < v.size() * MAXBITS) nbits = do_cnt();
    const word_t sz = nbits + anb;
    if (sz == nt) return;
    if (nt > sz) {
        if (nv > nt) nv = nt;
        if (nv > sz) {
            appendFill(1, nv - sz);
            if (nt > nv) appendFill(0, nt - nv);
        } else {
            appendFill(0, nt - sz);
        }
    } else {
        throw std::runtime_error("oracle adjustSize: truncation is not restated");
    }
}

// ---------------------------------------------------------------------------
// The four compressed-compressed loops.  Each step consumes the same number
// of words from both operands; the emitted tokens (append_counter,
// append_active, raw copies through copy_runs) decide the words that survive
// compress().

// src/bitvector.cpp:2419-2542
void BV::and_c2(const BV& rhs, BV& res) const {
    res.clear();
    if (v.size() == 1) {
        word_t w = v[0];
        if (w > HEADER1) { res.v = rhs.v; res.nbits = rhs.nbits; res.nset = rhs.nset; }
        else if (w > ALLONES) { res.v = v; res.nbits = nbits; res.nset = 0; }
        else { res.v.push_back(w & rhs.v[0]); res.nbits = nbits; }
    } else if (rhs.v.size() == 1) {
        word_t w = rhs.v[0];
        if (w > HEADER1) { res.v = v; res.nbits = nbits; res.nset = nset; }
        else if (w > ALLONES) { res.v = rhs.v; res.nbits = rhs.nbits; res.nset = 0; }
        else { res.v.push_back(w & v[0]); res.nbits = nbits; }
    } else if (v.size() > 1) {
        Run x, y;
        x.it = v.data(); y.it = rhs.v.data();
        const word_t* xend = v.data() + v.size();
        const word_t* yend = rhs.v.data() + rhs.v.size();
        while (x.it < xend) {
            if (x.nWords == 0) x.decode();
            if (y.nWords == 0) y.decode();
            if (x.isFill != 0) {
                if (y.isFill != 0) {
                    if (y.fillBit == 0) {
                        res.append_counter(0, y.nWords);
                        x -= y.nWords; y.nWords = 0; ++y.it;
                    } else if (y.nWords >= x.nWords) {
                        res.copy_runs(x, y.nWords);
                        y.it += (y.nWords == 0);
                    } else {
                        res.append_counter(x.fillBit, y.nWords);
                        x.nWords -= y.nWords; y.nWords = 0; ++y.it;
                    }
                } else if (x.fillBit == 0) {
                    res.append_counter(0, x.nWords);
                    y -= x.nWords; x.nWords = 0; ++x.it;
                } else {
                    res.copy_runs(y, x.nWords);
                    x.it += (x.nWords == 0);
                }
            } else if (y.isFill != 0) {
                if (y.fillBit == 0) {
                    res.append_counter(0, y.nWords);
                    x -= y.nWords; y.nWords = 0; ++y.it;
                } else {
                    res.copy_runs(x, y.nWords);
                    y.it += (y.nWords == 0);
                }
            } else {
                res.aval = *(x.it) & *(y.it);
                res.append_active();
                x.nWords = 0; y.nWords = 0; ++x.it; ++y.it;
            }
        }
        if (x.it != xend || y.it != yend) throw std::runtime_error("and_c2 internal error");
    }
    if (anb) { res.aval = aval & rhs.aval; res.anb = anb; }
}

// src/bitvector.cpp:2833-2978
void BV::or_c2(const BV& rhs, BV& res) const {
    res.clear();
    if (v.size() == 1) {
        word_t w = v[0];
        if (w > HEADER1) { res.v = v; res.nbits = nbits; res.nset = nbits; }
        else if (w > ALLONES) { res.v = rhs.v; res.nbits = rhs.nbits; res.nset = rhs.nset; }
        else { res.v.push_back(w | rhs.v[0]); res.nbits = nbits; }
    } else if (rhs.v.size() == 1) {
        word_t w = rhs.v[0];
        if (w > HEADER1) { res.v = rhs.v; res.nbits = rhs.nbits; res.nset = rhs.nbits; }
        else if (w > ALLONES) { res.v = v; res.nbits = nbits; res.nset = nset; }
        else { res.v.push_back(w | v[0]); res.nbits = nbits; }
    } else if (v.size() > 1) {
        Run x, y;
        x.it = v.data(); y.it = rhs.v.data();
        const word_t* xend = v.data() + v.size();
        const word_t* yend = rhs.v.data() + rhs.v.size();
        while (x.it < xend) {
            if (x.nWords == 0) x.decode();
            if (y.nWords == 0) y.decode();
            if (x.isFill != 0) {
                if (y.isFill != 0) {
                    if (y.fillBit != 0) {
                        res.append_counter(y.fillBit, y.nWords);
                        x -= y.nWords; y.nWords = 0; ++y.it;
                    } else if (y.nWords < x.nWords) {
                        res.append_counter(x.fillBit, y.nWords);
                        x.nWords -= y.nWords; y.nWords = 0; ++y.it;
                    } else {
                        res.copy_runs(x, y.nWords);
                        y.it += (y.nWords == 0);
                    }
                } else if (x.fillBit) {
                    res.append_counter(x.fillBit, x.nWords);
                    y -= x.nWords; x.nWords = 0; ++x.it;
                } else {
                    res.copy_runs(y, x.nWords);
                    x.it += (x.nWords == 0);
                }
            } else if (y.isFill != 0) {
                if (y.fillBit) {
                    res.append_counter(y.fillBit, y.nWords);
                    x -= y.nWords; y.nWords = 0; ++y.it;
                } else {
                    res.copy_runs(x, y.nWords);
                    y.it += (y.nWords == 0);
                }
            } else {
                res.aval = *x.it | *y.it;
                res.append_active();
                x.nWords = 0; y.nWords = 0; ++x.it; ++y.it;
            }
        }
        if (x.it != xend || y.it != yend) throw std::runtime_error("or_c2 internal error");
    }
    res.aval = aval | rhs.aval;
    res.anb = anb;
}

// src/bitvector.cpp:3281-3351
void BV::xor_c2(const BV& rhs, BV& res) const {
    res.clear();
    Run x, y;
    x.it = v.data(); y.it = rhs.v.data();
    const word_t* xend = v.data() + v.size();
    const word_t* yend = rhs.v.data() + rhs.v.size();
    while (x.it < xend) {
        if (x.nWords == 0) x.decode();
        if (y.nWords == 0) y.decode();
        if (x.isFill != 0) {
            if (y.isFill != 0 && y.nWords >= x.nWords) {
                if (y.fillBit == 0) res.copy_runs(x, y.nWords);
                else res.copy_runsn(x, y.nWords);
                y.it += (y.nWords == 0);
            } else if (x.fillBit == 0) {
                res.copy_runs(y, x.nWords);
                x.it += (x.nWords == 0);
            } else {
                res.copy_runsn(y, x.nWords);
                x.it += (x.nWords == 0);
            }
        } else if (y.isFill != 0) {
            if (y.fillBit == 0) res.copy_runs(x, y.nWords);
            else res.copy_runsn(x, y.nWords);
            y.it += (y.nWords == 0);
        } else {
            res.aval = *x.it ^ *y.it;
            res.append_active();
            x.nWords = 0; y.nWords = 0; ++x.it; ++y.it;
        }
    }
    if (x.it != xend || y.it != yend) throw std::runtime_error("xor_c2 internal error");
    res.aval = aval ^ rhs.aval;
    res.anb = anb;
}

// src/bitvector.cpp:3622-3745
void BV::minus_c2(const BV& rhs, BV& res) const {
    res.clear();
    if (v.size() == 1) {
        word_t w = v[0];
        if (w > HEADER1) {
            res.v.resize(rhs.v.size());
            for (size_t i = 0; i < rhs.v.size(); ++i)
                res.v[i] = rhs.v[i] > ALLONES ? (rhs.v[i] ^ FILLBIT) : (rhs.v[i] ^ ALLONES);
            res.nbits = rhs.nbits;
            if (rhs.nset > 0) res.nset = rhs.nbits - rhs.nset;
        } else if (w > ALLONES) {
            res.v = v; res.nbits = nbits; res.nset = 0;
        } else {
            res.v.push_back(w & ~rhs.v[0]); res.nbits = nbits;
        }
    } else if (rhs.v.size() == 1) {
        word_t w = rhs.v[0];
        if (w > HEADER1) res.appendFill(0, rhs.nbits);
        else if (w > ALLONES) { res.v = v; res.nbits = nbits; res.nset = nset; }
        else { res.v.push_back(w & v[0]); res.nbits = nbits; }
    } else if (v.size() > 1) {
        Run x, y;
        x.it = v.data(); y.it = rhs.v.data();
        const word_t* xend = v.data() + v.size();
        const word_t* yend = rhs.v.data() + rhs.v.size();
        while (x.it < xend) {
            if (x.nWords == 0) x.decode();
            if (y.nWords == 0) y.decode();
            if (x.isFill != 0) {
                if (y.isFill != 0 && y.nWords >= x.nWords) {
                    if (y.fillBit == 0) {
                        res.copy_runs(x, y.nWords);
                        y.it += (y.nWords == 0);
                    } else {
                        res.append_counter(0, y.nWords);
                        x -= y.nWords; y.nWords = 0; ++y.it;
                    }
                } else if (x.fillBit == 0) {
                    res.append_counter(0, x.nWords);
                    y -= x.nWords; x.nWords = 0; ++x.it;
                } else {
                    res.copy_runsn(y, x.nWords);
                    x.it += (x.nWords == 0);
                }
            } else if (y.isFill != 0) {
                if (y.fillBit == 0) {
                    res.copy_runs(x, y.nWords);
                    y.it += (y.nWords == 0);
                } else {
                    res.append_counter(0, y.nWords);
                    x -= y.nWords; y.nWords = 0; ++y.it;
                }
            } else {
                res.aval = *x.it & ~(*y.it);
                res.append_active();
                x.nWords = 0; y.nWords = 0; ++x.it; ++y.it;
            }
        }
        if (x.it != xend || y.it != yend) throw std::runtime_error("minus_c2 internal error");
    }
    res.aval = aval & ~(rhs.aval);
    res.anb = anb;
}

// src/bitvector.cpp:3354-3398 (rhs is literal-only)
void BV::xor_c1(const BV& rhs, BV& res) const {
    res.clear();
    size_t i1 = 0;
    for (size_t i0 = 0; i0 < v.size(); ++i0) {
        word_t w = v[i0];
        if (w > ALLONES) {
            word_t s0 = w & MAXCNT;
            res.nbits += s0 * MAXBITS;
            if (w >= HEADER1) for (word_t k = 0; k < s0; ++k, ++i1) res.v.push_back(rhs.v[i1] ^ ALLONES);
            else for (word_t k = 0; k < s0; ++k, ++i1) res.v.push_back(rhs.v[i1]);
        } else {
            res.aval = w ^ rhs.v[i1];
            res.append_active();
            ++i1;
        }
    }
    if (i1 != rhs.v.size()) throw std::runtime_error("xor_c1 internal error");
    res.aval = aval ^ rhs.aval;
    res.anb = anb;
}

// src/bitvector.cpp:3748-3818 (rhs is literal-only)
void BV::minus_c1(const BV& rhs, BV& res) const {
    res.clear();
    if (v.size() == 1) {
        word_t w = v[0];
        if (w > HEADER1) {
            res.v.resize(rhs.v.size());
            for (size_t i = 0; i < rhs.v.size(); ++i) res.v[i] = rhs.v[i] ^ ALLONES;
            res.nbits = rhs.nbits;
            if (rhs.nset > 0) res.nset = rhs.nbits - rhs.nset;
        } else if (w > ALLONES) {
            res.v = v; res.nbits = nbits; res.nset = 0;
        } else {
            res.v.push_back(w & ~rhs.v[0]); res.nbits = nbits;
        }
    } else if (v.size() > 1) {
        size_t i1 = 0;
        for (size_t i0 = 0; i0 < v.size(); ++i0) {
            word_t w = v[i0];
            if (w > ALLONES) {
                word_t s0 = w & MAXCNT;
                if (w >= HEADER1) {
                    for (word_t k = 0; k < s0; ++k, ++i1) res.v.push_back(rhs.v[i1] ^ ALLONES);
                    res.nbits += s0 * MAXBITS;
                } else {
                    i1 += s0;
                    if (s0 > 1) res.append_counter(0, s0);
                    else res.append_active();
                }
            } else {
                res.aval = w & ~rhs.v[i1];
                res.append_active();
                ++i1;
            }
        }
        if (i1 != rhs.v.size()) throw std::runtime_error("minus_c1 internal error");
    }
    res.aval = aval & ~(rhs.aval);
    res.anb = anb;
}

// ---------------------------------------------------------------------------
// src/bitvector.cpp:1194-1284 (&=), 1413-1483 (|=), 1557-1618 (^=),
// 1685-1762 (-=).  Returns the routine that produced the result.
static void active_and(BV& a, const BV& b) {
    if (a.anb == b.anb) a.aval &= b.aval;
    else if (a.anb > b.anb) a.aval &= (b.aval << (a.anb - b.anb));
    else { a.aval = (a.aval << (b.anb - a.anb)) & b.aval; a.anb = b.anb; }
}

Path BV::op_inplace(const BV& rhs, Op op) {
    if (size() > rhs.size()) {
        BV tmp(rhs);
        tmp.adjustSize(0, size());
        return op_inplace(tmp, op);
    } else if (size() < rhs.size()) {
        adjustSize(0, rhs.size());
    }
    const bool ca = (v.size() * MAXBITS == nbits && nbits > 0);
    const bool cb = (rhs.v.size() * MAXBITS == rhs.nbits && rhs.nbits > 0);
    const word_t dsum = (word_t)((v.size() + rhs.v.size()) * MAXBITS);

    // literal result: *_c0, *_d1, *_d2 (and the d1 after a swap)
    auto literal_result = [&](Path p) {
        std::vector<word_t> la = literal_words();
        std::vector<word_t> lb = rhs.literal_words();
        std::vector<word_t> out;
        literal_op(la, lb, op, out);
        word_t nb = (word_t)(out.size() * MAXBITS);
        v.swap(out);
        nbits = nb;
        nset = 0;
        aval = op_word(aval, rhs.aval, op);
        return p;
    };
    auto c2_result = [&](void (BV::*fn)(const BV&, BV&) const) {
        BV res;
        (this->*fn)(rhs, res);
        *this = res;
        return PATH_C2;
    };

    switch (op) {
    case OP_AND:
        if (ca) return literal_result(cb ? PATH_C0 : PATH_D1);
        if (cb) return literal_result(PATH_D1_SWAP);
        if (all0s() || rhs.all1s()) { active_and(*this, rhs); return PATH_SHORTCUT; }
        if (rhs.all0s() || all1s()) {
            nset = rhs.nset; v = rhs.v;
            active_and(*this, rhs);
            return PATH_SHORTCUT;
        }
        if (dsum >= rhs.nbits) return literal_result(PATH_D2);
        return c2_result(&BV::and_c2);
    case OP_OR:
        if (ca) return literal_result(cb ? PATH_C0 : PATH_D1);
        if (cb) return literal_result(PATH_D1_SWAP);
        if (all1s() || rhs.all0s()) { aval |= rhs.aval; return PATH_SHORTCUT; }
        if (all0s() || rhs.all1s()) {
            nset = rhs.nset; v = rhs.v; aval |= rhs.aval;
            return PATH_SHORTCUT;
        }
        if (dsum >= rhs.nbits) return literal_result(PATH_D2);
        return c2_result(&BV::or_c2);
    case OP_XOR:
        if (ca) return literal_result(cb ? PATH_C0 : PATH_D1);
        if (cb) {
            BV res; xor_c1(rhs, res); *this = res;
            return PATH_C1;
        }
        if (dsum >= rhs.nbits) return literal_result(PATH_D2);
        return c2_result(&BV::xor_c2);
    default: // OP_MINUS
        if (ca) return literal_result(cb ? PATH_C0 : PATH_D1);
        if (cb) {
            BV res; minus_c1(rhs, res); *this = res;
            return PATH_C1;
        }
        if (all0s() || rhs.all0s()) { aval &= ~(rhs.aval); return PATH_SHORTCUT; }
        if (rhs.all1s()) {
            nset = 0; nbits = 0; v.clear();
            aval &= ~(rhs.aval);
            append_counter(0, rhs.v[0] & MAXCNT);
            return PATH_SHORTCUT;
        }
        if (all1s()) {
            word_t tmp = aval;
            *this = rhs;
            flip();
            aval &= tmp;
            return PATH_SHORTCUT;
        }
        if (dsum >= rhs.nbits) return literal_result(PATH_D2);
        return c2_result(&BV::minus_c2);
    }
}

// Ascending positions of the set bits (src/bitvector.cpp:4524-4590 walks the
// same words; a literal word's bit 30 is its first position).
void BV::rids(std::vector<word_t>& out) const {
    word_t pos = 0;
    for (size_t i = 0; i < v.size(); ++i) {
        word_t w = v[i];
        if (w >= HEADER1) {
            word_t n = (w & MAXCNT) * MAXBITS;
            for (word_t k = 0; k < n; ++k) out.push_back(pos + k);
            pos += n;
        } else if (w >= HEADER0) {
            pos += (w & MAXCNT) * MAXBITS;
        } else {
            for (unsigned b = 0; b < MAXBITS; ++b)
                if ((w >> (SECONDBIT - b)) & 1U) out.push_back(pos + b);
            pos += MAXBITS;
        }
    }
    for (unsigned b = 0; b < anb; ++b)
        if ((aval >> (anb - 1 - b)) & 1U) out.push_back(pos + b);
}

} // namespace orc

namespace orc {

// ---------------------------------------------------------------------------
// Out-of-place operators.

// src/bitvector.cpp:2545-2610 (rhs is literal-only)
void BV::and_c1(const BV& rhs, BV& res) const {
    res.clear();
    if (v.size() == 1) {
        const word_t w = v[0];
        if (w > HEADER1) { res.v = rhs.v; res.nbits = rhs.nbits; res.nset = rhs.nset; }
        else if (w > ALLONES) { res.v = v; res.nbits = nbits; res.nset = 0; }
        else { res.v.push_back(w & rhs.v[0]); res.nbits = nbits; }
    } else if (v.size() > 1) {
        size_t i1 = 0;
        for (size_t i0 = 0; i0 < v.size(); ++i0) {
            const word_t w = v[i0];
            if (w > ALLONES) {
                const word_t s0 = w & MAXCNT;
                if (w < HEADER1) {
                    if (s0 > 1) res.append_counter(0, s0);
                    else res.append_active();
                    i1 += s0;
                } else {
                    for (word_t k = 0; k < s0; ++k) res.v.push_back(rhs.v[i1++]);
                    res.nbits += s0 * MAXBITS;
                }
            } else {
                res.aval = w & rhs.v[i1];
                res.append_active();
                ++i1;
            }
        }
        if (i1 != rhs.v.size()) throw std::runtime_error("and_c1 internal error");
    }
    res.aval = aval & rhs.aval;
    res.anb = anb;
}

// src/bitvector.cpp:2981-3050 (rhs is literal-only)
void BV::or_c1(const BV& rhs, BV& res) const {
    res.clear();
    if (v.size() == 1) {
        const word_t w = v[0];
        if (w > HEADER1) { res.v = v; res.nbits = nbits; res.nset = nbits; }
        else if (w > ALLONES) { res.v = rhs.v; res.nbits = rhs.nbits; res.nset = rhs.nset; }
        else { res.v.push_back(w | rhs.v[0]); res.nbits = nbits; }
    } else if (v.size() > 1) {
        size_t i1 = 0;
        for (size_t i0 = 0; i0 < v.size(); ++i0) {
            const word_t w = v[i0];
            if (w > ALLONES) {
                const word_t s0 = w & MAXCNT;
                if (w >= HEADER1) {
                    if (s0 > 1) res.append_counter(1, s0);
                    else { res.aval = ALLONES; res.append_active(); }
                    i1 += s0;
                } else {
                    for (word_t k = 0; k < s0; ++k) res.v.push_back(rhs.v[i1++]);
                    res.nbits += s0 * MAXBITS;
                }
            } else {
                res.aval = w | rhs.v[i1];
                res.append_active();
                ++i1;
            }
        }
        if (i1 != rhs.v.size()) throw std::runtime_error("or_c1 internal error");
    }
    res.aval = aval | rhs.aval;
    res.anb = anb;
}

// src/bitvector.cpp:3821-3869 (this is literal-only, rhs drives the loop)
void BV::minus_c1x(const BV& rhs, BV& res) const {
    res.clear();
    size_t i0 = 0;
    for (size_t i1 = 0; i1 < rhs.v.size(); ++i1) {
        const word_t w = rhs.v[i1];
        if (w > ALLONES) {
            const word_t s0 = w & MAXCNT;
            if (w >= HEADER1) {
                i0 += s0;
                if (s0 > 1) res.append_counter(0, s0);
                else res.append_active();
            } else {
                for (word_t k = 0; k < s0; ++k) res.v.push_back(v[i0++]);
                res.nbits += s0 * MAXBITS;
            }
        } else {
            res.aval = v[i0] & ~w;
            res.append_active();
            ++i0;
        }
    }
    if (i0 != v.size()) throw std::runtime_error("minus_c1x internal error");
    res.aval = aval & ~(rhs.aval);
    res.anb = anb;
}

// The operator's dispatch: operands of different sizes go through a copy of
// the smaller one, extended, and the in-place operator (with the copy on
// the left, except for MINUS); otherwise one routine writes a new vector.
BV BV::op_new(const BV& rhs, Op op, Path* path) const {
    BV res;
    Path p = PATH_NONE;
    if (size() > rhs.size()) {
        if (op == OP_MINUS) {
            BV tmp(rhs);
            tmp.adjustSize(0, size());
            res = *this;
            p = res.op_inplace(tmp, op);
        } else {
            res = rhs;
            res.adjustSize(0, size());
            p = res.op_inplace(*this, op);
        }
        if (path) *path = p;
        return res;
    } else if (size() < rhs.size()) {
        res = *this;
        res.adjustSize(0, rhs.size());
        p = res.op_inplace(rhs, op);
        if (path) *path = p;
        return res;
    }
    // MINUS tests literal-ness without nbits > 0 (src/bitvector.cpp:1788-1789)
    const bool ca = (v.size() * MAXBITS == nbits) && (nbits > 0 || op == OP_MINUS);
    const bool cb = (rhs.v.size() * MAXBITS == rhs.nbits) && (rhs.nbits > 0 || op == OP_MINUS);
    const bool d2 = (v.size() + rhs.v.size()) * MAXBITS > nbits;
    auto literal = [&](Path q) {
        std::vector<word_t> la = literal_words(), lb = rhs.literal_words();
        literal_op(la, lb, op, res.v);
        res.nbits = nbits;
        res.aval = op_word(aval, rhs.aval, op);
        res.anb = anb;
        p = q;
    };
    switch (op) {
    case OP_AND:
        if (ca && cb) literal(PATH_C0);
        else if (ca) { rhs.and_c1(*this, res); p = PATH_C1; }
        else if (cb) { and_c1(rhs, res); p = PATH_C1; }
        else if (all0s() || rhs.all1s()) { res = *this; res.aval &= rhs.aval; p = PATH_SHORTCUT; }
        else if (all1s() || rhs.all0s()) { res = rhs; res.aval &= aval; p = PATH_SHORTCUT; }
        else if (d2) literal(PATH_D2);
        else { and_c2(rhs, res); p = PATH_C2; }
        break;
    case OP_OR:
        if (ca && cb) literal(PATH_C0);
        else if (ca) { rhs.or_c1(*this, res); p = PATH_C1; }
        else if (cb) { or_c1(rhs, res); p = PATH_C1; }
        else if (all1s() || rhs.all0s()) { res = *this; res.aval |= rhs.aval; p = PATH_SHORTCUT; }
        else if (all0s() || rhs.all1s()) { res = rhs; res.aval |= aval; p = PATH_SHORTCUT; }
        else if (d2) literal(PATH_D2);
        else { or_c2(rhs, res); p = PATH_C2; }
        break;
    case OP_XOR:
        if (ca && cb) literal(PATH_C0);
        else if (ca) { rhs.xor_c1(*this, res); p = PATH_C1; }
        else if (cb) { xor_c1(rhs, res); p = PATH_C1; }
        else if (d2) literal(PATH_D2);
        else { xor_c2(rhs, res); p = PATH_C2; }
        break;
    default:
        if (ca && cb) literal(PATH_C0);
        else if (ca) { minus_c1x(rhs, res); p = PATH_C1; }
        else if (cb) { minus_c1(rhs, res); p = PATH_C1; }
        else if (all0s() || rhs.all0s()) { res = *this; res.aval &= ~rhs.aval; p = PATH_SHORTCUT; }
        else if (rhs.all1s()) {
            res.append_counter(0, rhs.v[0] & MAXCNT);
            res.anb = anb;
            res.aval = aval & ~rhs.aval;
            p = PATH_SHORTCUT;
        } else if (all1s()) { res = rhs; res.flip(); res.aval &= aval; p = PATH_SHORTCUT; }
        else if (d2) literal(PATH_D2);
        else { minus_c2(rhs, res); p = PATH_C2; }
        break;
    }
    if (path) *path = p;
    return res;
}

// ---------------------------------------------------------------------------
// count(mask)

word_t BV::count_mask(const BV& mask) const {
    word_t cnt = 0;
    const bool ca = (v.size() * MAXBITS == nbits && nbits > 0);
    const bool cb = (mask.v.size() * MAXBITS == mask.nbits && mask.nbits > 0);
    if (ca && cb) {
        for (size_t j = 0; j < v.size(); ++j) cnt += cnt_ones(v[j] & mask.v[j]);
        cnt += cnt_ones(aval & mask.aval);
    } else if (ca) {
        cnt = mask.count_c1(*this);
    } else if (cb) {
        cnt = count_c1(mask);
    } else if (all0s() || mask.all0s()) {
        cnt = cnt_ones(aval & mask.aval);
    } else if (all1s()) {
        if (mask.nset == 0 && !mask.v.empty()) mask.nbits = mask.do_cnt();
        cnt = mask.nset + cnt_ones(aval & mask.aval);
    } else if (mask.all1s()) {
        if (nbits == 0 && !v.empty()) nbits = do_cnt();
        cnt = nset + cnt_ones(aval & mask.aval);
    } else {
        cnt = count_c2(mask);
    }
    return cnt;
}

// src/bitvector.cpp:2317-2387
word_t BV::count_c2(const BV& mask) const {
    word_t cnt = cnt_ones(aval & mask.aval);
    Run x, y;
    x.it = v.data();
    y.it = mask.v.data();
    const word_t* xe = v.data() + v.size();
    const word_t* ye = mask.v.data() + mask.v.size();
    while (x.it < xe && y.it < ye) {
        if (x.nWords == 0) x.decode();
        if (y.nWords == 0) y.decode();
        if (x.isFill != 0) {
            if (y.isFill != 0) {
                if (x.fillBit != 0 && y.fillBit != 0) {
                    if (x.nWords <= y.nWords) {
                        cnt += MAXBITS * x.nWords;
                        y.nWords -= x.nWords; x.nWords = 0; ++x.it;
                        y.it += (y.nWords == 0);
                    } else {
                        cnt += MAXBITS * y.nWords;
                        x.nWords -= y.nWords; y.nWords = 0; ++y.it;
                    }
                } else if (x.nWords <= y.nWords) {
                    y.nWords -= x.nWords; y.it += (y.nWords == 0);
                    x.nWords = 0; ++x.it;
                } else {
                    x.nWords -= y.nWords; y.nWords = 0; ++y.it;
                }
            } else {
                if (x.fillBit != 0) cnt += cnt_ones(*y.it);
                ++y.it; --x.nWords; y.nWords = 0;
                x.it += (x.nWords == 0);
            }
        } else if (y.isFill != 0) {
            if (y.fillBit != 0) cnt += cnt_ones(*x.it);
            x.nWords = 0; --y.nWords; ++x.it;
            y.it += (y.nWords == 0);
        } else {
            cnt += cnt_ones(*x.it & *y.it);
            x.nWords = 0; y.nWords = 0; ++x.it; ++y.it;
        }
    }
    return cnt;
}

// src/bitvector.cpp:2390-2416 (mask is literal-only)
word_t BV::count_c1(const BV& mask) const {
    word_t cnt = cnt_ones(aval & mask.aval);
    size_t it = 0;
    for (size_t i = 0; i < v.size() && it < mask.v.size(); ++i) {
        const word_t w = v[i];
        if (w > ALLONES) {
            word_t n = w & MAXCNT;
            if (w > HEADER1) { while (n > 0) { cnt += cnt_ones(mask.v[it]); ++it; --n; } }
            else it += n;
        } else {
            cnt += cnt_ones(mask.v[it] & w);
            ++it;
        }
    }
    return cnt;
}

// ---------------------------------------------------------------------------
// erase

// src/bitvector.cpp:280-331
void BV::appendWord(word_t w) {
    const int cps = (int)(w >> MAXBITS);
    nset = 0;
    if (anb) {
        const word_t nb1 = anb;
        word_t nb2 = MAXBITS - anb;
        aval <<= nb2;
        if (cps != 0) {
            const int b2 = (w >= HEADER1);
            if (b2 != 0) aval |= (1U << nb2) - 1;
            append_active();
            nb2 = (w & MAXCNT) - 1;
            if (nb2 > 1) append_counter(b2, nb2);
            else if (nb2 == 1) { if (b2 != 0) aval = ALLONES; append_active(); }
            anb = nb1;
            aval = ((1U << nb1) - 1) * b2;
        } else {
            aval |= (w >> nb1);
            append_active();
            aval = (w & ((1U << nb1) - 1));
            anb = nb1;
        }
    } else if (cps != 0) {
        const int b2 = (w >= HEADER1);
        const word_t n = w & MAXCNT;
        if (n > 1) append_counter(b2, n);
        else if (n == 1) { if (b2) aval = ALLONES; append_active(); }
    } else {
        aval = w;
        append_active();
    }
}

namespace {
// The bit iterator of src/bitvector.cpp:4437-4518 (forward steps only):
// the code word holding a position, the offset in it, and the word's kind.
struct BitPos {
    const BV* bv;
    size_t it;              // index into m_vec; m_vec.size() = the active word
    word_t ind = 0;         // offset of the position in the word
    word_t nb = 0;          // bits in the word
    int compressed = 0, fillbit = 0;
    word_t literal = 0;     // literal value (the active word left-aligned)
    void decode() {
        const size_t n = bv->v.size();
        if (it < n) {
            const word_t w = bv->v[it];
            if (w > HEADER1) { fillbit = 1; compressed = 1; nb = (w & MAXCNT) * MAXBITS; }
            else if (w > HEADER0) { fillbit = 0; compressed = 1; nb = (w & MAXCNT) * MAXBITS; }
            else { compressed = 0; nb = MAXBITS; literal = w; }
        } else if (it == n) {
            compressed = 0; nb = bv->anb;
            literal = bv->aval << (MAXBITS - nb);
            it += (nb == 0);
        } else {
            it = n + 1; compressed = 0; nb = 0; literal = 0; fillbit = 0;
        }
        ind = 0;
    }
    void advance(word_t incr) {
        if (incr == 0) return;
        if (ind + incr < nb) { ind += incr; return; }
        long long rest = (long long)incr + ind - nb;
        while (rest >= 0 && it < bv->v.size()) {
            ++it;
            decode();
            if ((long long)nb > rest) { ind = (word_t)rest; rest = -1; }
            else rest -= nb;
        }
    }
};
} // namespace

// src/bitvector.cpp:1012-1081
// The size is counted first (do_cnt, with its collapse of an all-zero
// vector), as erase reads nbits as it stands.
void BV::erase(word_t i, word_t j) {
    size();
    if (i >= j) return;
    BV res;
    if (i > 0) {
        BitPos ip{this, 0};
        ip.decode();
        ip.advance(i);
        for (size_t c = 0; c < ip.it && c < v.size(); ++c) res.v.push_back(v[c]);
        res.nbits = i - ip.ind;
        if (ip.compressed) res.appendFill(ip.fillbit, ip.ind);
        else { res.aval = ip.literal >> (MAXBITS - ip.ind); res.anb = ip.ind; }
    }
    if (j < nbits) {
        BitPos iq{this, 0};
        iq.decode();
        iq.advance(j);
        if (iq.compressed) {
            res.appendFill(iq.fillbit, iq.nb - iq.ind);
        } else {
            for (long ii = (long)(iq.nb - iq.ind - 1); ii >= 0; --ii) res.push_bit(1 & (iq.literal >> ii));
        }
        for (size_t k = iq.it + 1; k < v.size(); ++k) res.appendWord(v[k]);
        for (long ii = (long)anb - 1; ii >= 0; --ii) res.push_bit((int)((aval >> ii) & 1));
    } else if (j < nbits + anb) {
        for (long ii = (long)(anb - j + nbits - 1); ii >= 0; --ii) res.push_bit((int)((aval >> ii) & 1));
    }
    *this = res;
}

} // namespace orc
