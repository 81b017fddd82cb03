// Host/device core of the WAH logical operators.
//
// This header is compiled twice: by hipcc into the gfx950 kernels and by the
// host compiler into the CPU self-test driver, so that the step logic can be
// exercised on the CPU at arbitrary partition points.
//
// Word format (FastBit 32-bit WAH, src/bitvector.cpp:18-31): bit 31 set marks
// a fill word, bit 30 is the fill value and bits 0-29 count 31-bit groups;
// a literal word holds 31 bits with the first bit in bit 30.
//
// Result model.  After compress() a result is a sequence of "positions"
// (31-bit groups).  Each position has a value and, when that value is all
// zeros or all ones, a representation F (it came from a fill token) or L (it
// came from a literal token).  compress() merges every maximal run of equal
// fill-valued positions: a run of two or more becomes one fill word; a run of
// exactly one stays a literal word when its representation is L and becomes a
// count-1 fill word when it is F (src/bitvector.cpp:378-469 never turns a
// count-1 fill back into a literal).  The representations are fixed by the
// step structure of the *_c2 loops (src/bitvector.cpp:2419, 2833, 3281, 3622)
// and of copy_runs / append_counter (src/bitvector.h:741-910); the c0/d1/d2/c1
// routines emit only L for single positions, so their compressed results are
// canonical.
//
// Parallel decomposition.  Every step of a *_c2 loop lies inside one code word
// of one operand, so a position where both operands start a code word (a
// "common boundary") always starts a step with both runs freshly decoded.
// The position space is cut at common boundaries; each piece is walked
// sequentially by one thread with exactly the reference's step logic, and the
// pieces' outputs are stitched by an associative run summary.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define WAH_HD __host__ __device__ __forceinline__
#else
#define WAH_HD inline
#endif

namespace wah {

typedef uint32_t word_t;
static constexpr word_t ALLONES = 0x7FFFFFFFu;
static constexpr word_t MAXCNT = 0x3FFFFFFFu;
static constexpr word_t FILLBIT = 0x40000000u;
static constexpr word_t HEADER0 = 0x80000000u;
static constexpr word_t HEADER1 = 0xC0000000u;

enum Op : int { AND = 0, OR = 1, XOR = 2, MINUS = 3 };

// Position classes used by the run summary.
enum Cls : uint32_t { C_NONE = 0, C_ZERO = 1, C_ONE = 2, C_MIXED = 3 };

WAH_HD word_t word_len(word_t w) { return w > ALLONES ? (w & MAXCNT) : 1u; }

WAH_HD word_t apply_op(word_t a, word_t b, int op) {
    return op == AND ? (a & b) : op == OR ? (a | b) : op == XOR ? (a ^ b) : (a & ~b);
}

// Value of the 31-bit group at a code word (fills expand to 0 / ALLONES).
WAH_HD word_t group_value(word_t w) {
    return w > ALLONES ? ((w >= HEADER1) ? ALLONES : 0u) : w;
}

WAH_HD uint32_t class_of_value(word_t v) {
    return v == 0 ? C_ZERO : (v == ALLONES ? C_ONE : C_MIXED);
}

// ---------------------------------------------------------------------------
// Run summary of a range of positions.  Associative; the empty range (cls
// fields C_NONE, count 0) is the identity.
//   count    words headed in the range, counting a leading run as one word
//            even when it continues a run from the left
//   first    class of the first position
//   uniform  the whole range is one run of ZERO or ONE positions
//   trail    class of the trailing open run (ZERO/ONE), MIXED when the range
//            ends with a mixed literal
//   tlen     length of the trailing open run
//   trep     representation of the trailing run when tlen == 1 (1 = F)
struct Summary {
    uint32_t count;
    uint32_t tlen;
    uint32_t first;
    uint32_t trail;
    uint32_t uniform;
    uint32_t trep;
};

WAH_HD Summary summary_empty() {
    Summary s; s.count = 0; s.tlen = 0; s.first = C_NONE; s.trail = C_NONE;
    s.uniform = 0; s.trep = 0; return s;
}

WAH_HD Summary combine(const Summary& l, const Summary& r) {
    if (l.first == C_NONE) return r;
    if (r.first == C_NONE) return l;
    const bool merge = (l.trail == r.first) && (r.first == C_ZERO || r.first == C_ONE);
    Summary s;
    s.count = l.count + r.count - (merge ? 1u : 0u);
    s.first = l.first;
    s.uniform = (l.uniform && r.uniform && merge) ? 1u : 0u;
    if (r.uniform && merge) {
        s.trail = l.trail; s.tlen = l.tlen + r.tlen; s.trep = 0;
    } else {
        s.trail = r.trail; s.tlen = r.tlen; s.trep = r.trep;
    }
    return s;
}

// 64-bit packing for the single-pass look-back between tiles:
// [63:62] status (1 aggregate, 2 inclusive prefix), [61:34] count,
// [33:6] tlen, [5:4] first, [3:2] trail, [1] uniform, [0] trep.
WAH_HD uint64_t pack(const Summary& s, uint64_t status) {
    return (status << 62) | ((uint64_t)(s.count & 0x0FFFFFFFu) << 34) |
           ((uint64_t)(s.tlen & 0x0FFFFFFFu) << 6) | ((uint64_t)(s.first & 3u) << 4) |
           ((uint64_t)(s.trail & 3u) << 2) | ((uint64_t)(s.uniform & 1u) << 1) |
           (uint64_t)(s.trep & 1u);
}

WAH_HD Summary unpack(uint64_t v) {
    Summary s;
    s.count = (uint32_t)((v >> 34) & 0x0FFFFFFFu);
    s.tlen = (uint32_t)((v >> 6) & 0x0FFFFFFFu);
    s.first = (uint32_t)((v >> 4) & 3u);
    s.trail = (uint32_t)((v >> 2) & 3u);
    s.uniform = (uint32_t)((v >> 1) & 1u);
    s.trep = (uint32_t)(v & 1u);
    return s;
}

WAH_HD uint32_t status_of(uint64_t v) { return (uint32_t)(v >> 62); }

// ---------------------------------------------------------------------------
// Token consumers.  A token is a span of positions with one class: a literal
// (len 1, any value) or a fill-valued span with a representation.

// Builds the context-free summary of a thread's range.
struct CountSink {
    Summary s;
    WAH_HD void init() { s = summary_empty(); }
    WAH_HD void tok(uint32_t cls, uint32_t len, uint32_t repF, word_t /*lit*/) {
        if (s.first == C_NONE) {
            s.first = cls; s.count = 1; s.trail = cls;
            s.tlen = (cls == C_MIXED) ? 0u : len;
            s.trep = repF; s.uniform = (cls != C_MIXED) ? 1u : 0u;
            return;
        }
        if (cls != C_MIXED && cls == s.trail) {
            s.tlen += len; s.trep = 0;
            return;
        }
        s.uniform = 0;
        s.count += 1;
        s.trail = cls;
        s.tlen = (cls == C_MIXED) ? 0u : len;
        s.trep = repF;
    }
};

// Writes the words of a thread's range.  `pre` is the exact summary of all
// positions before the range; the run open at the range start continues from
// it.  A fill-valued run is written when the next token closes it, at the
// slot of its head; the last open run is closed by finish() only at the end
// of the bitvector.  The word emitter W is called as w(slot, word, headpos).
template <class W>
struct WriteSink {
    W* w;
    uint32_t next;      // next free output slot
    uint32_t rcls;      // class of the open run (C_NONE if none)
    uint32_t rlen;      // its length so far
    uint32_t rrep;      // its representation if rlen == 1
    uint32_t pos;       // position of the next token
    WAH_HD void init(W* ww, const Summary& pre, uint32_t startpos) {
        w = ww; next = pre.count; pos = startpos;
        if (pre.first != C_NONE && (pre.trail == C_ZERO || pre.trail == C_ONE)) {
            rcls = pre.trail; rlen = pre.tlen; rrep = pre.trep;
        } else {
            rcls = C_NONE; rlen = 0; rrep = 0;
        }
    }
    WAH_HD void close_run() {
        if (rcls == C_ZERO || rcls == C_ONE) {
            const word_t b = (rcls == C_ONE) ? 1u : 0u;
            word_t word;
            if (rlen >= 2 || rrep) word = (b ? HEADER1 : HEADER0) | rlen;
            else word = b ? ALLONES : 0u;
            (*w)(next - 1, word, pos - rlen);
        }
        rcls = C_NONE; rlen = 0; rrep = 0;
    }
    WAH_HD void tok(uint32_t cls, uint32_t len, uint32_t repF, word_t lit) {
        if (cls != C_MIXED && cls == rcls) {
            rlen += len; rrep = 0; pos += len;
            return;
        }
        close_run();
        ++next;
        if (cls == C_MIXED) {
            (*w)(next - 1, lit, pos);
        } else {
            rcls = cls; rlen = len; rrep = repF;
        }
        pos += len;
    }
    WAH_HD void finish() { close_run(); }
};

// ---------------------------------------------------------------------------
// Sequential run state (src/bitvector.h:225-261) over an abstract word
// accessor R (R::at(index) returns the code word at that index).
template <class R>
struct Run {
    const R* src;
    uint32_t it;        // index of the current code word
    word_t nWords;      // remaining groups of the current code word
    int isFill;
    int fillBit;
    WAH_HD word_t cur() const { return src->at(it); }
    WAH_HD void decode() {
        const word_t w = cur();
        fillBit = (w > HEADER1);
        if (w > ALLONES) { nWords = (w & MAXCNT); isFill = 1; }
        else { nWords = 1; isFill = 0; }
    }
    WAH_HD void minus_eq(word_t len) {
        while (len > 0) {
            if (nWords == 0) decode();
            if (isFill != 0) {
                if (nWords > len) { nWords -= len; len = 0; }
                else if (nWords == len) { nWords = 0; len = 0; ++it; }
                else { len -= nWords; ++it; nWords = 0; }
            } else { --len; ++it; nWords = 0; }
        }
    }
};

// Emits the token of one group-valued span (class from the value).
template <class S>
WAH_HD void emit_fill(S& sink, int b, word_t k, uint32_t repF) {
    sink.tok(b ? C_ONE : C_ZERO, k, repF, 0u);
}

template <class S>
WAH_HD void emit_lit(S& sink, word_t v, uint32_t canon) {
    (void)canon;
    const uint32_t c = class_of_value(v);
    sink.tok(c, 1u, 0u, v);
}

// copy_runs / copy_runsn (src/bitvector.h:838-910) as tokens: the first code
// word is appended through append_active or append_counter, the following
// whole code words are pushed raw while they fit into nw.
template <class R, class S>
WAH_HD void copy_runs_tok(S& sink, Run<R>& it, word_t& nw, bool neg, uint32_t canon) {
    const uint32_t F = canon ? 0u : 1u;
    if (it.isFill == 0) {
        const word_t v = neg ? (ALLONES ^ it.cur()) : it.cur();
        emit_lit(sink, v, canon);
        --nw;
    } else if (it.nWords > 1) {
        emit_fill(sink, neg ? !it.fillBit : it.fillBit, it.nWords, F);
        nw -= it.nWords;
    } else if (it.nWords == 1) {
        emit_fill(sink, neg ? !it.fillBit : it.fillBit, 1u, 0u);
        --nw;
    }
    ++it.it;
    it.nWords = 0;
    while (nw > 0) {
        it.decode();
        if (nw >= it.nWords) {
            const word_t w = it.cur();
            if (it.isFill) emit_fill(sink, neg ? !it.fillBit : it.fillBit, it.nWords, F);
            else emit_lit(sink, neg ? (ALLONES ^ w) : w, canon);
            nw -= it.nWords;
            it.nWords = 0;
            ++it.it;
        } else {
            break;
        }
    }
}

// One step of the *_c2 loop of `op`; returns the groups consumed.
// canon != 0 forces every token's representation to L (the canonical result
// of the c0/d1/d2/c1 routines).
template <class R, class S>
WAH_HD word_t c2_step(S& sink, Run<R>& x, Run<R>& y, int op, uint32_t canon) {
    const uint32_t F = canon ? 0u : 1u;
    if (x.nWords == 0) x.decode();
    if (y.nWords == 0) y.decode();
    word_t nw;
    if (op == AND) {                                   // src/bitvector.cpp:2460-2516
        if (x.isFill) {
            if (y.isFill) {
                if (y.fillBit == 0) {
                    nw = y.nWords; emit_fill(sink, 0, nw, F);
                    x.minus_eq(nw); y.nWords = 0; ++y.it; return nw;
                } else if (y.nWords >= x.nWords) {
                    nw = y.nWords; word_t r = y.nWords;
                    copy_runs_tok(sink, x, r, false, canon);
                    y.nWords = r; y.it += (r == 0); return nw - r;
                } else {
                    nw = y.nWords; emit_fill(sink, x.fillBit, nw, F);
                    x.nWords -= nw; y.nWords = 0; ++y.it; return nw;
                }
            } else if (x.fillBit == 0) {
                nw = x.nWords; emit_fill(sink, 0, nw, F);
                y.minus_eq(nw); x.nWords = 0; ++x.it; return nw;
            } else {
                nw = x.nWords; word_t r = x.nWords;
                copy_runs_tok(sink, y, r, false, canon);
                x.nWords = r; x.it += (r == 0); return nw - r;
            }
        } else if (y.isFill) {
            if (y.fillBit == 0) {
                nw = y.nWords; emit_fill(sink, 0, nw, F);
                x.minus_eq(nw); y.nWords = 0; ++y.it; return nw;
            } else {
                nw = y.nWords; word_t r = y.nWords;
                copy_runs_tok(sink, x, r, false, canon);
                y.nWords = r; y.it += (r == 0); return nw - r;
            }
        }
    } else if (op == OR) {                             // src/bitvector.cpp:2874-2934
        if (x.isFill) {
            if (y.isFill) {
                if (y.fillBit != 0) {
                    nw = y.nWords; emit_fill(sink, 1, nw, F);
                    x.minus_eq(nw); y.nWords = 0; ++y.it; return nw;
                } else if (y.nWords < x.nWords) {
                    nw = y.nWords; emit_fill(sink, x.fillBit, nw, F);
                    x.nWords -= nw; y.nWords = 0; ++y.it; return nw;
                } else {
                    nw = y.nWords; word_t r = y.nWords;
                    copy_runs_tok(sink, x, r, false, canon);
                    y.nWords = r; y.it += (r == 0); return nw - r;
                }
            } else if (x.fillBit) {
                nw = x.nWords; emit_fill(sink, 1, nw, F);
                y.minus_eq(nw); x.nWords = 0; ++x.it; return nw;
            } else {
                nw = x.nWords; word_t r = x.nWords;
                copy_runs_tok(sink, y, r, false, canon);
                x.nWords = r; x.it += (r == 0); return nw - r;
            }
        } else if (y.isFill) {
            if (y.fillBit) {
                nw = y.nWords; emit_fill(sink, 1, nw, F);
                x.minus_eq(nw); y.nWords = 0; ++y.it; return nw;
            } else {
                nw = y.nWords; word_t r = y.nWords;
                copy_runs_tok(sink, x, r, false, canon);
                y.nWords = r; y.it += (r == 0); return nw - r;
            }
        }
    } else if (op == XOR) {                            // src/bitvector.cpp:3287-3325
        if (x.isFill) {
            if (y.isFill && y.nWords >= x.nWords) {
                nw = y.nWords; word_t r = y.nWords;
                copy_runs_tok(sink, x, r, y.fillBit != 0, canon);
                y.nWords = r; y.it += (r == 0); return nw - r;
            } else {
                nw = x.nWords; word_t r = x.nWords;
                copy_runs_tok(sink, y, r, x.fillBit != 0, canon);
                x.nWords = r; x.it += (r == 0); return nw - r;
            }
        } else if (y.isFill) {
            nw = y.nWords; word_t r = y.nWords;
            copy_runs_tok(sink, x, r, y.fillBit != 0, canon);
            y.nWords = r; y.it += (r == 0); return nw - r;
        }
    } else {                                           // MINUS, src/bitvector.cpp:3668-3717
        if (x.isFill) {
            if (y.isFill && y.nWords >= x.nWords) {
                if (y.fillBit == 0) {
                    nw = y.nWords; word_t r = y.nWords;
                    copy_runs_tok(sink, x, r, false, canon);
                    y.nWords = r; y.it += (r == 0); return nw - r;
                } else {
                    nw = y.nWords; emit_fill(sink, 0, nw, F);
                    x.minus_eq(nw); y.nWords = 0; ++y.it; return nw;
                }
            } else if (x.fillBit == 0) {
                nw = x.nWords; emit_fill(sink, 0, nw, F);
                y.minus_eq(nw); x.nWords = 0; ++x.it; return nw;
            } else {
                nw = x.nWords; word_t r = x.nWords;
                copy_runs_tok(sink, y, r, true, canon);
                x.nWords = r; x.it += (r == 0); return nw - r;
            }
        } else if (y.isFill) {
            if (y.fillBit == 0) {
                nw = y.nWords; word_t r = y.nWords;
                copy_runs_tok(sink, x, r, false, canon);
                y.nWords = r; y.it += (r == 0); return nw - r;
            } else {
                nw = y.nWords; emit_fill(sink, 0, nw, F);
                x.minus_eq(nw); y.nWords = 0; ++y.it; return nw;
            }
        }
    }
    // both literal: append_active of the combined word
    emit_lit(sink, apply_op(x.cur(), y.cur(), op), canon);
    x.nWords = 0; y.nWords = 0; ++x.it; ++y.it;
    return 1;
}

// Walks the c2 loop from a common boundary (code word ia of A and ib of B,
// both starting at position p0) until position p1 (another common boundary).
template <class R, class S>
WAH_HD void walk(S& sink, const R* ra, const R* rb, uint32_t ia, uint32_t ib,
                 uint32_t p0, uint32_t p1, int op, uint32_t canon) {
    Run<R> x, y;
    x.src = ra; x.it = ia; x.nWords = 0; x.isFill = 0; x.fillBit = 0;
    y.src = rb; y.it = ib; y.nWords = 0; y.isFill = 0; y.fillBit = 0;
    uint32_t p = p0;
    while (p < p1) p += c2_step(sink, x, y, op, canon);
}

// From code word ia of A starting at position pa and code word ib of B
// starting at pb, advance to the first position >= max(pa,pb) where both
// operands start a code word.  Stops at `end` (the total group count), which
// is always a common boundary.
template <class R>
WAH_HD uint32_t find_common(const R* ra, const R* rb, uint32_t& ia, uint32_t& pa,
                            uint32_t& ib, uint32_t& pb, uint32_t end) {
    while (pa != pb) {
        if (pa < pb) { pa += word_len(ra->at(ia)); ++ia; }
        else { pb += word_len(rb->at(ib)); ++ib; }
        if (pa >= end && pb >= end) break;
    }
    return pa;
}

struct ArrayReader {
    const word_t* p;
    WAH_HD word_t at(uint32_t i) const { return p[i]; }
};

} // namespace wah
