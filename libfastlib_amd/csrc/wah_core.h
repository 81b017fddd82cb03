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

// Builds the context-free summary of a thread's range.  Every field is
// assigned once per token from selects (stores chosen by branch would be
// merged into one store through a chosen field address, which keeps the
// summary in memory).
struct CountSink {
    Summary s;
    WAH_HD void init() { s = summary_empty(); }
    // merge == 0: a word pushed raw (never merged into the word before it,
    // though later tokens may merge into it); its class counts as MIXED
    // for the merge test of the range's first token
    WAH_HD void tok(uint32_t cls, uint32_t len, uint32_t repF, word_t /*lit*/, uint32_t merge) {
        const bool empty = s.first == C_NONE;
        const bool extend = !empty && merge && cls != C_MIXED && cls == s.trail;   // the open run goes on
        const uint32_t nlen = cls == C_MIXED ? 0u : len;
        Summary n;
        n.first = empty ? (merge ? cls : (uint32_t)C_MIXED) : s.first;
        n.count = empty ? 1u : (extend ? s.count : s.count + 1u);
        n.trail = extend ? s.trail : cls;
        n.tlen = extend ? s.tlen + len : nlen;
        n.trep = extend ? 0u : repF;
        n.uniform = empty ? (cls != C_MIXED ? 1u : 0u) : (extend ? s.uniform : 0u);
        s = n;
    }
};

// Writes the words of a thread's range.  `pre` is the exact summary of all
// positions before the range; the run open at the range start continues from
// it.  A fill-valued run is written when the next token closes it, at the
// slot of its head; the last open run is closed by finish() only at the end
// of the bitvector.  The word emitter W is called as w(slot, word, headpos).
template <class W>
struct WriteSink {
    W w;                // the word emitter (held by value)
    uint32_t next;      // next free output slot
    uint32_t rcls;      // class of the open run (C_NONE if none)
    uint32_t rlen;      // its length so far
    uint32_t rrep;      // its representation if rlen == 1
    uint32_t pos;       // position of the next token
    WAH_HD void init(const W* ww, const Summary& pre, uint32_t startpos) {
        w = *ww; next = pre.count; pos = startpos;
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
            w(next - 1, word, pos - rlen);
        }
        rcls = C_NONE; rlen = 0; rrep = 0;
    }
    WAH_HD void tok(uint32_t cls, uint32_t len, uint32_t repF, word_t lit, uint32_t merge) {
        const bool extend = merge && cls != C_MIXED && cls == rcls;    // the open run goes on
        const uint32_t olen = rlen;
        if (!extend) {
            close_run();
            ++next;
            if (cls == C_MIXED) w(next - 1, lit, pos);
        }
        // fields assigned once, from selects (see CountSink)
        rcls = extend ? cls : (cls == C_MIXED ? (uint32_t)C_NONE : cls);
        rlen = extend ? olen + len : (cls == C_MIXED ? 0u : len);
        rrep = extend || cls == C_MIXED ? 0u : repF;
        pos += len;
    }
    WAH_HD void finish() { close_run(); }
};

// ---------------------------------------------------------------------------
// Sequential run state (src/bitvector.h:225-261) over an abstract word
// accessor R (R::at(index) returns the code word at that index).
template <class R>
struct Run {
    R src;              // the words (held by value: runs are swapped by value)
    uint32_t it;        // index of the current code word
    word_t nWords;      // remaining groups of the current code word
    int isFill;
    int fillBit;
    WAH_HD word_t cur() const { return src.at(it); }
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

// Field-wise choice between two runs (readers provide R::pick): plain
// value selects, so that choosing a run never takes its address.
template <class R>
WAH_HD Run<R> pick(bool c, const Run<R>& a, const Run<R>& b) {
    Run<R> r;
    r.src = R::pick(c, a.src, b.src);
    r.it = c ? a.it : b.it;
    r.nWords = c ? a.nWords : b.nWords;
    r.isFill = c ? a.isFill : b.isFill;
    r.fillBit = c ? a.fillBit : b.fillBit;
    return r;
}

// Emits the token of one group-valued span (class from the value).
// merge: the token goes through append_counter / append_active (it merges
// into the word before it when the classes match); 0 for a word pushed raw.
template <class S>
WAH_HD void emit_fill(S& sink, int b, word_t k, uint32_t repF, uint32_t merge = 1u) {
    sink.tok(b ? C_ONE : C_ZERO, k, repF, 0u, merge);
}

template <class S>
WAH_HD void emit_lit(S& sink, word_t v, uint32_t merge = 1u) {
    const uint32_t c = class_of_value(v);
    sink.tok(c, 1u, 0u, v, merge);
}

// Mode bits of the step walk: MODE_CANON forces every representation to L
// (the canonical results of the c0/d1/d2/c1 routines); MODE_RAW emits the
// *_c2 loop's own words, not compressed: raw pushes of copy_runs do not
// merge into the word before them, and walks start only where both
// operands start a code word.
constexpr uint32_t MODE_CANON = 1u, MODE_RAW = 2u;

// copy_runs / copy_runsn (src/bitvector.h:838-910) as tokens: the first code
// word is appended through append_active or append_counter, the following
// whole code words are pushed raw while they fit into nw.
template <class R, class S>
WAH_HD void copy_runs_tok(S& sink, Run<R>& it, word_t& nw, bool neg, uint32_t mode) {
    const uint32_t F = (mode & MODE_CANON) ? 0u : 1u;
    const uint32_t mr = (mode & MODE_RAW) ? 0u : 1u;      // raw pushes merge only when compressed
    if (it.isFill == 0) {
        const word_t v = neg ? (ALLONES ^ it.cur()) : it.cur();
        emit_lit(sink, v);
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
            if (it.isFill) emit_fill(sink, neg ? !it.fillBit : it.fillBit, it.nWords, F, mr);
            else emit_lit(sink, neg ? (ALLONES ^ w) : w, mr);
            nw -= it.nWords;
            it.nWords = 0;
            ++it.it;
        } else {
            break;
        }
    }
}

// Steps of the *_c2 loops consume a span chosen by the branch; a step here
// consumes at most `lim` groups (lim >= 1), so that a walk ends exactly at a
// synchronisation point inside a span.  Cutting a span only ends a token
// early; the next walk's tokens continue the same run and merge with it.
template <class R>
WAH_HD void take(Run<R>& r, word_t n) {          // n <= r.nWords groups of the decoded word
    if (n == r.nWords) { r.nWords = 0; ++r.it; }
    else r.nWords -= n;
}

// One step of the *_c2 loop of `op`; returns the groups consumed.
// mode & MODE_CANON forces every token's representation to L (the canonical result
// of the c0/d1/d2/c1 routines).  The branch of the reference's loop is first
// reduced to an action and the operand it is anchored on (f, the other o);
// f and o are chosen from x and y field by field and written back the same
// way, so that each action runs on fixed operands (a run chosen through a
// reference would keep both runs in memory instead of registers).
enum StepAct : int { S_LIT = 0, S_FILL_SKIP = 1, S_COPY_IN = 2, S_FILL_SHORT = 3 };

template <class R, class S>
WAH_HD word_t c2_step(S& sink, Run<R>& x, Run<R>& y, int op, uint32_t mode, word_t lim = 0xFFFFFFFFu) {
    const uint32_t F = (mode & MODE_CANON) ? 0u : 1u;
    if (x.nWords == 0) x.decode();
    if (y.nWords == 0) y.decode();
    // act on (f, o): f = x, or f = y when `swp`
    int act = S_LIT, bit = 0;
    bool swp = false, neg = false;
    if (op == AND) {                                   // src/bitvector.cpp:2460-2516
        if (x.isFill) {
            if (y.isFill) {
                if (y.fillBit == 0) { act = S_FILL_SKIP; swp = true; bit = 0; }
                else if (y.nWords >= x.nWords) { act = S_COPY_IN; swp = true; }
                else act = S_FILL_SHORT;
            } else if (x.fillBit == 0) { act = S_FILL_SKIP; bit = 0; }
            else act = S_COPY_IN;
        } else if (y.isFill) {
            if (y.fillBit == 0) { act = S_FILL_SKIP; swp = true; bit = 0; }
            else { act = S_COPY_IN; swp = true; }
        }
    } else if (op == OR) {                             // src/bitvector.cpp:2874-2934
        if (x.isFill) {
            if (y.isFill) {
                if (y.fillBit != 0) { act = S_FILL_SKIP; swp = true; bit = 1; }
                else if (y.nWords < x.nWords) act = S_FILL_SHORT;
                else { act = S_COPY_IN; swp = true; }
            } else if (x.fillBit) { act = S_FILL_SKIP; bit = 1; }
            else act = S_COPY_IN;
        } else if (y.isFill) {
            if (y.fillBit) { act = S_FILL_SKIP; swp = true; bit = 1; }
            else { act = S_COPY_IN; swp = true; }
        }
    } else if (op == XOR) {                            // src/bitvector.cpp:3287-3325
        if (x.isFill) {
            if (y.isFill && y.nWords >= x.nWords) { act = S_COPY_IN; swp = true; neg = y.fillBit != 0; }
            else { act = S_COPY_IN; neg = x.fillBit != 0; }
        } else if (y.isFill) {
            act = S_COPY_IN; swp = true; neg = y.fillBit != 0;
        }
    } else {                                           // MINUS, src/bitvector.cpp:3668-3717
        if (x.isFill) {
            if (y.isFill && y.nWords >= x.nWords) {
                if (y.fillBit == 0) { act = S_COPY_IN; swp = true; }
                else { act = S_FILL_SKIP; swp = true; bit = 0; }
            } else if (x.fillBit == 0) { act = S_FILL_SKIP; bit = 0; }
            else { act = S_COPY_IN; neg = true; }
        } else if (y.isFill) {
            if (y.fillBit == 0) { act = S_COPY_IN; swp = true; }
            else { act = S_FILL_SKIP; swp = true; bit = 0; }
        }
    }
    if (act == S_LIT) {
        // both literal: append_active of the combined word
        emit_lit(sink, apply_op(x.cur(), y.cur(), op));
        x.nWords = 0; y.nWords = 0; ++x.it; ++y.it;
        return 1;
    }
    Run<R> f = pick(swp, y, x), o = pick(swp, x, y);
    word_t used;
    if (act == S_FILL_SKIP) {
        // a fill of f spans the step, o's words are skipped
        used = f.nWords < lim ? f.nWords : lim;
        emit_fill(sink, bit, used, F);
        o.minus_eq(used);
        take(f, used);
    } else if (act == S_COPY_IN) {
        // the words of o are copied within the fill of f
        const word_t nw = f.nWords < lim ? f.nWords : lim;
        word_t r = nw;
        copy_runs_tok(sink, o, r, neg, mode);
        used = nw - r;
        f.nWords -= used;
        f.it += (f.nWords == 0);
    } else {
        // both fills, o is the shorter and ends first: the span has f's bit
        used = o.nWords < lim ? o.nWords : lim;
        emit_fill(sink, f.fillBit, used, F);
        f.nWords -= used;
        take(o, used);
    }
    x = pick(swp, o, f);
    y = pick(swp, f, o);
    return used;
}

// Walks the c2 loop over [p0, p1) between two synchronisation points (see
// find_sync).  ia / ib are the code words of A and B at p0 and pa / pb
// their start positions: an operand whose word starts before p0 is inside
// a fill there, with pa + len - p0 groups of it left.
template <class R, class S>
WAH_HD void walk(S& sink, const R* ra, const R* rb, uint32_t ia, uint32_t pa, uint32_t ib, uint32_t pb,
                 uint32_t p0, uint32_t p1, int op, uint32_t mode) {
    Run<R> x, y;
    x.src = *ra; x.it = ia; x.nWords = 0; x.isFill = 0; x.fillBit = 0;
    y.src = *rb; y.it = ib; y.nWords = 0; y.isFill = 0; y.fillBit = 0;
    if (p0 < p1) {
        if (pa < p0) { x.decode(); x.nWords = pa + x.nWords - p0; }
        if (pb < p0) { y.decode(); y.nWords = pb + y.nWords - p0; }
    }
    uint32_t p = p0;
    while (p < p1) p += c2_step(sink, x, y, op, mode, p1 - p);
}

// Synchronisation points of the c2 loop: positions where the loop's state
// is fixed by the position alone, so that a walk may start there -- both
// operands start a code word, or one starts a code word (not a fill of one
// group) strictly inside a fill of the other.  (A fill of one group at the
// start of a copied span could be appended instead of pushed raw.)  From
// code word ia of A starting at pa and code word ib of B starting at pb,
// returns the first such position >= max(pa, pb), visiting every word start
// in order (so that the search is monotone in its start), with ia / pa and
// ib / pb the words holding it and their starts.  `end` (the total group
// count, or any synchronisation point) bounds the search.
template <class R>
WAH_HD uint32_t find_sync(const R* ra, const R* rb, uint32_t& ia, uint32_t& pa, uint32_t& ib, uint32_t& pb,
                          uint32_t end, bool raw = false) {
    uint32_t q = pa > pb ? pa : pb;
    if (q >= end) return end;
    // the words holding q
    while (pa < q) { const uint32_t l = word_len(ra->at(ia)); if (pa + l > q) break; pa += l; ++ia; }
    while (pb < q) { const uint32_t l = word_len(rb->at(ib)); if (pb + l > q) break; pb += l; ++ib; }
    for (;;) {
        if (q >= end) return end;
        const word_t wa = ra->at(ia), wb = rb->at(ib);
        const uint32_t la = word_len(wa), lb = word_len(wb);
        if (pa == q && pb == q) return q;
        // raw words: a copied span may not restart at a word start (its
        // first word would be appended, not pushed raw)
        if (!raw && pa == q && !(wa > ALLONES && la == 1u)) return q;
        if (!raw && pb == q && !(wb > ALLONES && lb == 1u)) return q;
        const uint32_t na = pa + la, nb = pb + lb;
        q = na < nb ? na : nb;
        if (na == q) { pa = na; ++ia; }
        if (nb == q) { pb = nb; ++ib; }
    }
}

// The *_c2 loop of op writes the canonical words of its result except for
// a result group g that is a run of its own (all 0 or all 1, its neighbours
// of another class): that group's word is a count-1 fill when its token is
// a fill and a literal when it is a literal.  c2_rep_at returns true for a
// fill.  (ia, pa) / (ib, pb): the code words of A / B holding g and their
// first groups.  The walk starts at the last synchronisation point at or
// before g (find_sync's rule, taken backwards: the later of the two word
// starts, unless that word is a fill of one group starting inside the
// other operand's fill) and runs the step logic up to g.  The token that
// covers g has length 1, since g is a run of its own, so cutting the walk
// after g leaves it as the full loop emits it.
struct RepSink {
    uint32_t pos, g, rep;
    WAH_HD void tok(uint32_t /*cls*/, uint32_t len, uint32_t repF, word_t /*lit*/, uint32_t /*merge*/) {
        if (pos <= g && g - pos < len) rep = repF;
        pos += len;
    }
};

template <class R>
WAH_HD bool c2_rep_at(const R* ra, const R* rb, uint32_t ia, uint32_t pa, uint32_t ib, uint32_t pb, uint32_t g,
                      int op) {
    for (;;) {
        if (pa == pb) break;
        if (pa > pb) {
            const word_t w = ra->at(ia);
            if (!(w > ALLONES && word_len(w) == 1u)) break;
            --ia;
            pa -= word_len(ra->at(ia));
        } else {
            const word_t w = rb->at(ib);
            if (!(w > ALLONES && word_len(w) == 1u)) break;
            --ib;
            pb -= word_len(rb->at(ib));
        }
    }
    const uint32_t s = pa > pb ? pa : pb;
    RepSink sink{s, g, 0u};
    walk(sink, ra, rb, ia, pa, ib, pb, s, g + 1u, op, 0u);
    return sink.rep != 0u;
}

struct ArrayReader {
    const word_t* p;
    WAH_HD word_t at(uint32_t i) const { return p[i]; }
    static WAH_HD ArrayReader pick(bool c, const ArrayReader& a, const ArrayReader& b) {
        return ArrayReader{c ? a.p : b.p};
    }
};

} // namespace wah
