#include "runtime.h"
#include "../../include/fastbit_hip.h"
#include <map>
#include <mutex>
#include <vector>

namespace fbh_rt {

static int g_device = -1;
static hipStream_t g_stream = nullptr;
static hipStream_t g_side[NSIDE] = {};
static std::mutex g_mu;
static std::multimap<size_t, void*> g_free;        // rounded size -> buffer
static std::map<void*, size_t> g_size;             // live buffer -> rounded size
static thread_local std::string t_err;

// Scratch of the allocation-free entry points: a file-scope device array of
// NOALLOC_WORDS words that every GPU the library runs on holds until the
// process ends.  In host code its name stands for a host-side shadow, not
// for device memory, so its device address is looked up once per device
// (at initialisation, with that device current) and only that address is
// handed to kernels.
static __device__ __align__(256) uint32_t noalloc_scratch[NOALLOC_WORDS];
constexpr int MAX_DEVICES = 64;
static uint32_t* g_noalloc[MAX_DEVICES] = {};

std::recursive_mutex& api_mutex() {
    static std::recursive_mutex m;
    return m;
}

void set_error(const std::string& s) { t_err = s; }
const char* last_error() { return t_err.c_str(); }

int check(hipError_t e, const char* what) {
    if (e == hipSuccess) return FBH_OK;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return FBH_ERR_DEVICE;
}

int ensure_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_stream && (device < 0 || device == g_device)) return FBH_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error("no HIP device available");
        return FBH_ERR_DEVICE;
    }
    if (device < 0) device = 0;
    if (device >= n) { set_error("device index out of range"); return FBH_ERR_ARG; }
    if ((e = hipSetDevice(device)) != hipSuccess) return check(e, "hipSetDevice");
    if (g_stream) { (void)hipStreamDestroy(g_stream); g_stream = nullptr; }
    for (hipStream_t& t : g_side)
        if (t) { (void)hipStreamDestroy(t); t = nullptr; }
    if ((e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking)) != hipSuccess)
        return check(e, "hipStreamCreate");
    for (hipStream_t& t : g_side)
        if ((e = hipStreamCreateWithFlags(&t, hipStreamNonBlocking)) != hipSuccess)
            return check(e, "hipStreamCreate");
    if (device < MAX_DEVICES && !g_noalloc[device]) {
        void* p = nullptr;
        if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch))) != hipSuccess)
            return check(e, "hipGetSymbolAddress");
        g_noalloc[device] = (uint32_t*)p;
    }
    g_device = device;
    return FBH_OK;
}

uint32_t* noalloc_scratch_ptr() {
    return (g_device >= 0 && g_device < MAX_DEVICES) ? g_noalloc[g_device] : nullptr;
}

hipStream_t stream() { return g_stream; }
hipStream_t side_stream(int k) { return g_side[k % NSIDE]; }

static size_t round_size(size_t b) {
    if (b < 256) return 256;
    size_t r = 256;
    while (r < b) r <<= 1;
    return r;
}

void* dev_alloc(size_t bytes) {
    const size_t r = round_size(bytes);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_free.find(r);
        if (it != g_free.end()) {
            void* p = it->second;
            g_free.erase(it);
            g_size[p] = r;
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, r);
    if (e != hipSuccess) {
        // release cached buffers and retry once
        std::lock_guard<std::mutex> lk(g_mu);
        (void)hipStreamSynchronize(g_stream);
        for (auto& kv : g_free) (void)hipFree(kv.second);
        g_free.clear();
        e = hipMalloc(&p, r);
        if (e != hipSuccess) { set_error("hipMalloc failed"); return nullptr; }
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_size[p] = r;
    return p;
}

void dev_free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_size.find(p);
    if (it == g_size.end()) return;
    g_free.emplace(it->second, p);
    g_size.erase(it);
}

void* Scratch::get(size_t need) {
    if (need <= bytes && p) return p;
    if (p) dev_free(p);
    size_t b = need < 4096 ? 4096 : need;
    b += b / 4;
    p = dev_alloc(b);
    bytes = p ? b : 0;
    return p;
}

Scratch& scratch(int slot) {
    static Scratch s[8];
    return s[slot & 7];
}

} // namespace fbh_rt
