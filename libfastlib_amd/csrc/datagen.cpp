// Deterministic synthetic workloads (host C++, libfbgen.so).
//
// Every generator uses a seeded xoshiro256** stream and only IEEE basic
// operations (+ - * / sqrt, frexp/ldexp), with its own log/exp built from
// them, so that the same seed gives bit-identical data on every x86-64 host
// that runs this binary.  Build with -ffp-contract=off.
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <vector>
#include <algorithm>

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) { for (int i = 0; i < 4; ++i) s[i] = splitmix(seed); }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

const double LN2 = 0.69314718055994530942;

// log(x), x > 0: m*2^e with m in [sqrt(1/2), sqrt(2)), atanh series.
double det_log(double x) {
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) { m *= 2.0; e -= 1; }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double t = s, sum = s;
    for (int k = 3; k <= 41; k += 2) { t *= s2; sum += t / (double)k; }
    return 2.0 * sum + (double)e * LN2;
}

// exp(x): x = k ln2 + r, |r| <= ln2/2, Taylor series.
double det_exp(double x) {
    const double kf = floor(x / LN2 + 0.5);
    const double r = x - kf * LN2;
    double t = 1.0, sum = 1.0;
    for (int k = 1; k <= 30; ++k) { t *= r / (double)k; sum += t; }
    return ldexp(sum, (int)kf);
}

// Geometric run length (>= 1) with success probability p.
uint32_t geom(Rng& g, double p, double log1mp) {
    if (p >= 1.0) return 1u;
    double u = g.uniform();
    if (u <= 0.0) u = 1e-300;
    const double k = floor(det_log(u) / log1mp);
    if (k > 4.0e9) return 0xFFFFFFFFu;
    return 1u + (uint32_t)k;
}

double std_normal(Rng& g) {
    for (;;) {
        const double u = 2.0 * g.uniform() - 1.0;
        const double v = 2.0 * g.uniform() - 1.0;
        const double s = u * u + v * v;
        if (s > 0.0 && s < 1.0) return u * sqrt(-2.0 * det_log(s) / s);
    }
}

constexpr uint32_t ALLONES = 0x7FFFFFFFu;
constexpr uint32_t HEADER0 = 0x80000000u;
constexpr uint32_t HEADER1 = 0xC0000000u;

// Canonical WAH (serialized) of literal groups + active bits.
void encode(const std::vector<uint32_t>& lit, uint32_t aval, uint32_t anb,
            std::vector<uint32_t>& out) {
    out.clear();
    size_t i = 0;
    while (i < lit.size()) {
        const uint32_t w = lit[i];
        if (w == 0 || w == ALLONES) {
            size_t j = i + 1;
            while (j < lit.size() && lit[j] == w) ++j;
            const size_t n = j - i;
            if (n >= 2) out.push_back((w ? HEADER1 : HEADER0) | (uint32_t)n);
            else out.push_back(w);
            i = j;
        } else {
            out.push_back(w);
            ++i;
        }
    }
    if (anb) out.push_back(aval);
    out.push_back(anb);
}

uint32_t* hand_out(const std::vector<uint32_t>& v, size_t* n) {
    uint32_t* p = (uint32_t*)malloc(sizeof(uint32_t) * (v.size() ? v.size() : 1));
    if (p && !v.empty()) memcpy(p, v.data(), sizeof(uint32_t) * v.size());
    *n = v.size();
    return p;
}

} // namespace

extern "C" {

void fbg_free(void* p) { free(p); }

// Two-state Markov bit source: set-bit density d, mean 1-run length L
// (p10 = 1/L, p01 = d p10 / (1-d)), initial state Bernoulli(d).  Returns the
// canonical serialized WAH of nbits bits.
int fbg_markov_wah(uint64_t seed, uint32_t nbits, double d, double L, uint32_t** out, size_t* n) {
    Rng g(seed);
    const double p10 = 1.0 / L;
    double p01 = d * p10 / (1.0 - d);
    if (p01 > 1.0) p01 = 1.0;
    const double l10 = p10 < 1.0 ? det_log(1.0 - p10) : 0.0;
    const double l01 = p01 < 1.0 ? det_log(1.0 - p01) : 0.0;
    int state = g.uniform() < d ? 1 : 0;
    const uint32_t groups = nbits / 31u;
    const uint32_t anb = nbits - groups * 31u;
    std::vector<uint32_t> lit(groups, 0u);
    uint32_t aval = 0;
    uint64_t pos = 0;
    while (pos < nbits) {
        uint64_t len = state ? geom(g, p10, l10) : geom(g, p01, l01);
        if (pos + len > nbits) len = nbits - pos;
        if (state) {
            // set bits [pos, pos+len): word-granular fill where possible
            uint64_t b = pos;
            const uint64_t e = pos + len;
            while (b < e) {
                const uint64_t wi = b / 31u;
                if (wi >= groups) {
                    for (; b < e; ++b) aval |= 1u << (anb - 1 - (uint32_t)(b - (uint64_t)groups * 31u));
                    break;
                }
                const uint32_t off = (uint32_t)(b - wi * 31u);
                const uint64_t wend = (wi + 1) * 31u;
                const uint32_t cnt = (uint32_t)((e < wend ? e : wend) - b);
                // bits off .. off+cnt-1 of the group; bit 0 is the MSB (bit 30)
                const uint32_t mask = (cnt == 31u) ? ALLONES
                                    : (((1u << cnt) - 1u) << (31u - off - cnt));
                lit[wi] |= mask;
                b += cnt;
            }
        }
        pos += len;
        state ^= 1;
    }
    std::vector<uint32_t> s;
    encode(lit, aval, anb, s);
    *out = hand_out(s, n);
    return *out ? 0 : -1;
}

// Uniform integers in [lo, hi).
int fbg_uniform_i32(uint64_t seed, uint32_t n, int32_t lo, int32_t hi, int32_t* out) {
    Rng g(seed);
    const uint64_t span = (uint64_t)((int64_t)hi - (int64_t)lo);
    for (uint32_t i = 0; i < n; ++i) out[i] = (int32_t)((int64_t)lo + (int64_t)g.below(span));
    return 0;
}

// Standard normal float32 values.
int fbg_normal_f32(uint64_t seed, uint32_t n, float* out) {
    Rng g(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = (float)std_normal(g);
    return 0;
}

// Zipf(s) integers on {1..K}: P(k) ~ k^-s (CDF table, binary search).
int fbg_zipf_i32(uint64_t seed, uint32_t n, double s, uint32_t K, int32_t* out) {
    Rng g(seed);
    std::vector<double> cdf(K);
    double acc = 0.0;
    for (uint32_t k = 1; k <= K; ++k) { acc += det_exp(-s * det_log((double)k)); cdf[k - 1] = acc; }
    for (uint32_t i = 0; i < n; ++i) {
        const double u = g.uniform() * acc;
        const uint32_t k = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        out[i] = (int32_t)((k < K ? k : K - 1) + 1);
    }
    return 0;
}

// Float column drawn from a pool of `pool` distinct values: a fraction
// `wnorm` of the pool (and of the rows) are N(0,1) draws, the rest is the
// Zipf(s) support {1, 2, ...}; normal rows pick a pool normal uniformly,
// Zipf rows pick support value k with probability ~ k^-s.  Writes the
// column and, if pool_out != NULL, the pool's normal values.
int fbg_float_mix(uint64_t seed, uint32_t n, uint32_t pool, double wnorm, double s, float* out,
                  float* pool_out) {
    Rng g(seed);
    const uint32_t nn = (uint32_t)(pool * wnorm);
    const uint32_t nz = pool - nn;
    std::vector<float> normals;
    normals.reserve(nn);
    // distinct normal values (reject the rare duplicates and Zipf integers)
    std::vector<float> seen;
    while (normals.size() < nn) {
        const float v = (float)std_normal(g);
        normals.push_back(v);
    }
    std::sort(normals.begin(), normals.end());
    normals.erase(std::unique(normals.begin(), normals.end()), normals.end());
    while (normals.size() < nn) {           // top up after de-duplication
        const float v = (float)std_normal(g);
        if (!std::binary_search(normals.begin(), normals.end(), v)) {
            normals.insert(std::upper_bound(normals.begin(), normals.end(), v), v);
        }
    }
    if (pool_out) memcpy(pool_out, normals.data(), sizeof(float) * nn);
    std::vector<double> cdf(nz);
    double acc = 0.0;
    for (uint32_t k = 1; k <= nz; ++k) { acc += det_exp(-s * det_log((double)k)); cdf[k - 1] = acc; }
    for (uint32_t i = 0; i < n; ++i) {
        if (g.uniform() < wnorm) {
            out[i] = normals[g.below(nn)];
        } else {
            const double u = g.uniform() * acc;
            uint32_t k = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
            if (k >= nz) k = nz - 1;
            out[i] = (float)(k + 1);
        }
    }
    return 0;
}

// Sorted runs: run lengths Geometric(p), run values increasing by 1..3.
int fbg_sorted_runs_i32(uint64_t seed, uint32_t n, double p, int32_t* out) {
    Rng g(seed);
    const double lp = det_log(1.0 - p);
    int32_t v = 0;
    uint32_t i = 0;
    while (i < n) {
        uint32_t len = geom(g, p, lp);
        if (len > n - i) len = n - i;
        for (uint32_t k = 0; k < len; ++k) out[i + k] = v;
        i += len;
        v += 1 + (int32_t)g.below(3);
    }
    return 0;
}

// Uniform doubles in [0,1) (query endpoint draws).
int fbg_uniform_f64(uint64_t seed, uint32_t n, double* out) {
    Rng g(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = g.uniform();
    return 0;
}

// Uniform integers in [lo, hi] (inclusive), int64 output.
int fbg_uniform_incl_i64(uint64_t seed, uint32_t n, int64_t lo, int64_t hi, int64_t* out) {
    Rng g(seed);
    const uint64_t span = (uint64_t)(hi - lo) + 1u;
    for (uint32_t i = 0; i < n; ++i) out[i] = lo + (int64_t)g.below(span);
    return 0;
}

} // extern "C"
