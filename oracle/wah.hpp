// CPU oracle: restatement of FastBit's 32-bit WAH bitvector (ibis::bitvector).
//
// TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library in
// libfastlib_amd/; only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline may load it.  The product path never calls it.
//
// Every routine cites the reference function it restates (paths relative to
// the FastBit 2.0.3 source tree).  The routines whose emitted words decide the
// bit-exact result after compress() (the *_c2 and *_c1 loops, copy_runs,
// append_active, append_counter, compress, do_cnt) follow the reference's step
// structure one to one; the routines that write a plain literal array
// (*_c0, *_d1, *_d2) are restated by what they compute, an op over the
// decompressed words, because their result words are fixed by the bits alone.
#pragma once
#include <cstdint>
#include <cstddef>
#include <vector>
#include <stdexcept>

namespace orc {

typedef uint32_t word_t;
// src/bitvector.cpp:18-31
static const unsigned MAXBITS = 31;
static const unsigned SECONDBIT = 30;
static const word_t ALLONES = 0x7FFFFFFFu;
static const word_t MAXCNT = 0x3FFFFFFFu;
static const word_t FILLBIT = 0x40000000u;
static const word_t HEADER0 = 0x80000000u;
static const word_t HEADER1 = 0xC0000000u;

enum Op { OP_AND = 0, OP_OR = 1, OP_XOR = 2, OP_MINUS = 3 };

// Which reference routine produced the result of an in-place op.
enum Path {
    PATH_NONE = 0, PATH_C0 = 1, PATH_D1 = 2, PATH_D1_SWAP = 3, PATH_C1 = 4,
    PATH_D2 = 5, PATH_C2 = 6, PATH_SHORTCUT = 7
};

struct BV;

// src/bitvector.h:225-261 (struct run)
struct Run {
    int isFill = 0;
    int fillBit = 0;
    word_t nWords = 0;
    const word_t* it = nullptr;
    void decode() {
        fillBit = (*it > HEADER1);
        if (*it > ALLONES) { nWords = (*it & MAXCNT); isFill = 1; }
        else { nWords = 1; isFill = 0; }
    }
    void operator--() {
        if (nWords > 1) { --nWords; } else { ++it; nWords = 0; }
    }
    void operator-=(word_t len) {
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

// src/bitvector.h:62-318 (class ibis::bitvector, 32-bit words)
struct BV {
    mutable word_t nbits = 0;     // bits in m_vec (0 = not yet counted)
    mutable word_t nset = 0;      // ones in m_vec (0 = not yet counted)
    word_t aval = 0;              // active word value (right aligned)
    word_t anb = 0;               // active word bit count
    mutable std::vector<word_t> v;  // m_vec

    // --- construction / serialization
    void clear() { nset = 0; nbits = 0; v.clear(); aval = 0; anb = 0; }
    static BV from_serial(const word_t* a, size_t n);   // src/bitvector.cpp:54-105
    std::vector<word_t> serial() const;                  // src/bitvector.cpp:2283-2313
    void set(int val, word_t n);                         // src/bitvector.cpp:256-276

    // --- append primitives (src/bitvector.h:699-831)
    void append_active();
    void append_counter(int val, word_t cnt);
    void push_bit(int b);                                // operator+=(int)
    void appendFill(int val, word_t n);
    void setBit(word_t ind, int val);                    // append path of src/bitvector.cpp:621-685

    // --- copy primitives used by the *_c2 loops (src/bitvector.h:838-910)
    void copy_runs(Run& it, word_t& nw);
    void copy_runsn(Run& it, word_t& nw);

    // --- queries
    word_t do_cnt() const;                               // src/bitvector.cpp:583-608
    word_t size() const { return ((nbits ? nbits : (nbits = do_cnt())) + anb); }
    word_t cnt() const;                                  // src/bitvector.h:652-656
    bool all0s() const;                                  // src/bitvector.h:592-603
    bool all1s() const;                                  // src/bitvector.h:607-613
    static word_t cnt_ones(word_t w) { return (word_t)__builtin_popcount(w); }

    // --- transforms
    void compress();                                     // src/bitvector.cpp:378-469
    void decompress();                                   // src/bitvector.cpp:474-512
    std::vector<word_t> literal_words() const;           // src/bitvector.cpp:515-541
    void flip();                                         // src/bitvector.cpp:1128-1174
    void adjustSize(word_t nv, word_t nt);               // src/bitvector.cpp:4086-4113 (growth)

    // --- in-place logical operators with the reference's routine dispatch
    Path op_inplace(const BV& rhs, Op op);               // src/bitvector.cpp:1194-1762

    // --- the compressed-compressed loops
    void and_c2(const BV& rhs, BV& res) const;           // src/bitvector.cpp:2419-2542
    void or_c2(const BV& rhs, BV& res) const;            // src/bitvector.cpp:2833-2978
    void xor_c2(const BV& rhs, BV& res) const;           // src/bitvector.cpp:3281-3351
    void minus_c2(const BV& rhs, BV& res) const;         // src/bitvector.cpp:3622-3745
    void xor_c1(const BV& rhs, BV& res) const;           // src/bitvector.cpp:3354-3398
    void minus_c1(const BV& rhs, BV& res) const;         // src/bitvector.cpp:3748-3818

    // --- rid extraction (src/bitvector.cpp:4524-4590, ascending positions)
    void rids(std::vector<word_t>& out) const;

    // --- out-of-place operators: a new bitvector, not compressed
    // (src/bitvector.cpp:1292-1407 &, 1488-1551 |, 1624-1679 ^, 1768-1838 -)
    BV op_new(const BV& rhs, Op op, Path* path) const;
    void and_c1(const BV& rhs, BV& res) const;           // src/bitvector.cpp:2545-2610
    void or_c1(const BV& rhs, BV& res) const;            // src/bitvector.cpp:2981-3050
    void minus_c1x(const BV& rhs, BV& res) const;        // src/bitvector.cpp:3821-3869

    // --- ones of this AND mask without the result (src/bitvector.cpp:1088-1124)
    word_t count_mask(const BV& mask) const;
    word_t count_c1(const BV& mask) const;               // src/bitvector.cpp:2390-2416
    word_t count_c2(const BV& mask) const;               // src/bitvector.cpp:2317-2387

    // --- erase bits [i, j) (src/bitvector.cpp:1012-1081) and the word
    // append it uses (src/bitvector.cpp:280-331)
    void erase(word_t i, word_t j);
    void appendWord(word_t w);
};

// Literal-array result of one op, as *_c0 / *_d1 / *_d2 produce it.
void literal_op(const std::vector<word_t>& a, const std::vector<word_t>& b,
                Op op, std::vector<word_t>& out);
word_t op_word(word_t a, word_t b, Op op);

} // namespace orc
