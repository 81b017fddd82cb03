// Library runtime: device selection, the library stream, a stream-ordered
// caching allocator for HBM buffers, and the thread-local error string.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <mutex>
#include <string>

namespace fbh_rt {

// Select the device and create the stream once; returns 0 or an FBH error.
int ensure_init(int device = -1);
hipStream_t stream();
// NSIDE more streams (default priority) for overlapping independent
// launches of one call; the call joins them back into stream() before it
// returns, so buffer reuse stays ordered on stream().
constexpr int NSIDE = 3;
hipStream_t side_stream(int k);

// Device address of the allocation-free entry points' scratch array on the
// current device (NOALLOC_WORDS words, taken once per device by
// ensure_init; nullptr before it).  Every call on a GPU shares it.
constexpr uint32_t NOALLOC_PARTS = 65536;             // partial counts (tiles or scan blocks)
constexpr uint32_t NOALLOC_WORDS = NOALLOC_PARTS + 64; // + the count
uint32_t* noalloc_scratch_ptr();

// Caching allocator.  Buffers are handed back to a free list keyed by the
// rounded size and reused by later requests; all library work runs on one
// stream, so reuse is ordered after the work that used the buffer.
void* dev_alloc(size_t bytes);
void dev_free(void* p);

// Library-wide lock.  Every C-ABI entry point holds it for its whole call:
// the library stream, the scratch slots and the look-back status tables
// are shared by all handles, so two host threads working on different
// handles must not interleave their launches and read-backs.  Recursive
// because entry points call one another.
std::recursive_mutex& api_mutex();
#define FBH_API_LOCK() std::lock_guard<std::recursive_mutex> fbh_api_guard_(fbh_rt::api_mutex())

void set_error(const std::string& s);
const char* last_error();

// Checks a HIP status; on failure records it and returns FBH_ERR_DEVICE.
int check(hipError_t e, const char* what);

// A growable device scratch region owned by the library (per process).
struct Scratch {
    void* p = nullptr;
    size_t bytes = 0;
    void* get(size_t need);
};
Scratch& scratch(int slot);

} // namespace fbh_rt
