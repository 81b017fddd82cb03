// Host launchers of the gfx950 kernels (defined in *.hip next to the
// kernels they launch).  All launches go to the caller's stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbk {
typedef uint32_t word_t;
constexpr uint32_t CHUNK_WORDS = 32;

hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s);
hipError_t launch_zero_u64(uint64_t* p, uint32_t n, hipStream_t s);
uint32_t scan_tmp_words(uint32_t n);
hipError_t launch_scan_excl(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp,
                            hipStream_t s);
hipError_t launch_chunk_index(const word_t* w, uint32_t n, uint32_t* ck, uint32_t* tmp,
                              hipStream_t s);
uint32_t op_tiles(uint32_t na, uint32_t nb);
// status entries launch_op needs for nt = op_tiles(na, nb) tiles
uint32_t op_status_words(uint32_t nt);
hipError_t launch_op(const word_t* a, const uint32_t* cka, uint32_t na,
                     const word_t* b, const uint32_t* ckb, uint32_t nb,
                     uint32_t groups, int op, int mode, word_t* out, uint32_t* outck,
                     uint64_t* status, uint32_t* ctl, hipStream_t s);
hipError_t launch_edge_words(const word_t* out, uint32_t* ctl, hipStream_t s);
uint32_t tok_tiles(uint32_t n);
hipError_t launch_tok(const word_t* a, const word_t* b, uint32_t n, uint32_t groups, int op,
                      int mode, word_t* out, uint32_t* outck, uint64_t* status, uint32_t* ctl,
                      hipStream_t s);
// out2[0] = ones, out2[1] = groups; out2 needs cnt_scratch_words() words.
hipError_t launch_cnt(const word_t* w, uint32_t n, uint32_t* out2, hipStream_t s);
uint32_t cnt_scratch_words();
hipError_t launch_decompress(const word_t* w, const uint32_t* ck, uint32_t n, uint32_t groups,
                             word_t* out, int neg, hipStream_t s);
hipError_t launch_flip(word_t* w, uint32_t n, hipStream_t s);
hipError_t launch_rids_lit(const word_t* lit, uint32_t n, uint32_t* out, uint32_t cap,
                           uint32_t* tmp, hipStream_t s);
// Canonical WAH of a OP b (b == nullptr: of a alone) group by group; an
// operand with lit != 0 holds one literal word per group.  status needs
// canon_status(groups) entries; ctl 8 words; ts 2 * (canon_tiles(groups) + 1)
// entries (tile starts of the compressed operands).
// After a canonical op kernel wrote the canonical words of a OP b (both
// operands compressed, with chunk indexes) to out / outck with their count
// in ctl[1] (at most maxwords): the *_c2 routines' words, every literal
// 0 / all-ones word that the routine writes as a count-1 fill rewritten.
hipError_t launch_c2_fix(const word_t* a, const uint32_t* cka, uint32_t na, const word_t* b, const uint32_t* ckb,
                         uint32_t nb, word_t* out, const uint32_t* outck, const uint32_t* ctl, uint32_t maxwords,
                         int op, hipStream_t s);
uint32_t canon_tiles(uint32_t groups);
uint32_t canon_status(uint32_t groups);
hipError_t launch_canon(const word_t* a, const uint32_t* cka, uint32_t na, int alit,
                        const word_t* b, const uint32_t* ckb, uint32_t nb, int blit,
                        uint32_t groups, int op, word_t* out, uint32_t* outck, uint64_t* status,
                        uint32_t* ctl, uint2* ts, hipStream_t s);
// count(mask): part[k] (count_and_blocks(groups) entries) = ones of a AND b
// over the k-th tile of groups; lit != 0: the operand is one literal word
// per group (no chunk index needed).
uint32_t count_and_blocks(uint32_t groups);
hipError_t launch_count_and(const word_t* a, const uint32_t* cka, uint32_t na, int alit, const word_t* b,
                            const uint32_t* ckb, uint32_t nb, int blit, uint32_t groups, uint32_t* part,
                            hipStream_t s);
// out[i] = a[i] OP b[i]
hipError_t launch_lit_op(const word_t* a, const word_t* b, uint32_t n, int op, word_t* out, hipStream_t s);
// *_c1 routines: V[g] = value of result group g, H[g] = 1 where a result
// word starts (x compressed with its chunk index, y literal).
enum : int { C1_AND = 0, C1_OR = 1, C1_XOR = 2, C1_MINUS = 3, C1_MINUSX = 4 };
hipError_t launch_c1_heads(const word_t* x, const uint32_t* ckx, uint32_t nx, const word_t* y, uint32_t groups,
                           int rk, word_t* V, uint32_t* H, hipStream_t s);
// Result words from the word starts H, their exclusive scan O and the
// values V of n groups; nw = number of words; HP needs nw + 1 words.
hipError_t launch_head_words(const word_t* V, const uint32_t* H, const uint32_t* O, uint32_t n, uint32_t nw,
                             uint32_t* HP, word_t* out, hipStream_t s);
// out[0] / out[1] = index and start group of the code word holding group G
hipError_t launch_word_at(const word_t* w, const uint32_t* ck, uint32_t n, uint32_t G, uint32_t* out,
                          hipStream_t s);
// V[q] = group q of bits [p0, p0 + ind) ++ bits [j, j + nS - ind) of the
// literal groups L (the active bits left-aligned in one more group, then a
// zero word).
hipError_t launch_shift_groups(const word_t* L, uint32_t p0, uint32_t ind, uint32_t j, uint32_t nS, word_t* V,
                               hipStream_t s);
} // namespace fbk
