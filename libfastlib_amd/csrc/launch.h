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
hipError_t launch_op(const word_t* a, const uint32_t* cka, uint32_t na,
                     const word_t* b, const uint32_t* ckb, uint32_t nb,
                     uint32_t groups, int op, int canon, word_t* out, uint32_t* outck,
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
uint32_t canon_tiles(uint32_t groups);
uint32_t canon_status(uint32_t groups);
hipError_t launch_canon(const word_t* a, const uint32_t* cka, uint32_t na, int alit,
                        const word_t* b, const uint32_t* ckb, uint32_t nb, int blit,
                        uint32_t groups, int op, word_t* out, uint32_t* outck, uint64_t* status,
                        uint32_t* ctl, uint2* ts, hipStream_t s);
} // namespace fbk
