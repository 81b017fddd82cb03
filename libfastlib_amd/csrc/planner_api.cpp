// C ABI of the learned plan choice (include/fastbit_hip.h): an FP32 MLP
// cost model trained and evaluated by the kernels of mlp_kernels.hip.  The
// reference decides plans with fixed rules (column::evaluateRange,
// src/column.cpp:6102-6111; bin::sumBins, src/index.cpp:6640-6783); those
// rules stay the fallback.
#include "../../include/fastbit_hip.h"
#include "runtime.h"
#include "mlp_dev.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace fbh_rt;

static_assert(fbm::NPARAM == FBH_PLANNER_NPARAM, "parameter count");
static_assert(fbm::NF == FBH_NFEATURES && fbm::NO == FBH_NPLANS, "shapes");

namespace {

int fail(int code, const char* msg) { set_error(msg); return code; }

// splitmix64 (Steele, Lea, Flood 2014): seeded weights and shuffles.
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0, 1)
    uint32_t below(uint32_t n) { return (uint32_t)(unit() * n); }
};

// A device array owned by the planner.
template <class T>
struct DArr {
    T* p = nullptr;
    size_t n = 0;
    ~DArr() { dev_free(p); }
    int ensure(size_t k) {
        if (k <= n && p) return FBH_OK;
        dev_free(p);
        p = (T*)dev_alloc(sizeof(T) * (k ? k : 1));
        n = p ? k : 0;
        return p ? FBH_OK : fail(FBH_ERR_NOMEM, "device allocation failed");
    }
    int put(const T* h, size_t k) {
        int rc = ensure(k);
        if (!rc && k) rc = check(hipMemcpyAsync(p, h, sizeof(T) * k, hipMemcpyHostToDevice, stream()), "upload");
        return rc;
    }
    int get(T* h, size_t k) const {
        int rc = k ? check(hipMemcpyAsync(h, p, sizeof(T) * k, hipMemcpyDeviceToHost, stream()), "download") : 0;
        if (!rc) rc = check(hipStreamSynchronize(stream()), "sync");
        return rc;
    }
};

} // namespace

struct fbh_planner {
    DArr<float> params, m, v, grad;
    DArr<float> mu, sd, lo, hi;
    // per-call buffers
    DArr<float> x, y, mask, xn, h1, h2, out, d1, d2, d3, loss;
    DArr<uint32_t> idx, nfall;
    DArr<int32_t> fb, plan;
    float hmu[fbm::NF], hsd[fbm::NF];
    bool trained = false;
    uint64_t t = 0;          // Adam steps taken

    int reserve(uint32_t n) {
        int rc;
        if ((rc = xn.ensure((size_t)n * fbm::NF)) || (rc = h1.ensure((size_t)n * fbm::H1)) ||
            (rc = h2.ensure((size_t)n * fbm::H2)) || (rc = out.ensure((size_t)n * fbm::NO)) ||
            (rc = d1.ensure((size_t)n * fbm::H1)) || (rc = d2.ensure((size_t)n * fbm::H2)) ||
            (rc = d3.ensure((size_t)n * fbm::NO)))
            return rc;
        return FBH_OK;
    }
    int put_norm() {
        int rc = mu.put(hmu, fbm::NF);
        if (!rc) rc = sd.put(hsd, fbm::NF);
        return rc;
    }
    // forward + deltas of the samples idx (null: 0..n-1) of the uploaded
    // x / y / mask; loss parts at lp
    int fwd_delta(const uint32_t* didx, uint32_t n, float inv_m, float* lp) {
        int rc = check(fbm::launch_fwd(params.p, x.p, didx, n, mu.p, sd.p, xn.p, h1.p, h2.p, out.p, stream()), "fwd");
        if (!rc) rc = check(fbm::launch_delta(params.p, didx, n, y.p, mask.p, inv_m, h1.p, h2.p, out.p, d1.p, d2.p,
                                              d3.p, lp, stream()), "delta");
        return rc;
    }
    int adam(float lr) {
        ++t;
        const double c1 = 1.0 - std::pow(0.9, (double)t), c2 = 1.0 - std::pow(0.999, (double)t);
        return check(fbm::launch_adam(params.p, grad.p, m.p, v.p, lr, 0.9f, 0.999f, 1e-8f, (float)c1, (float)c2,
                                      stream()), "adam");
    }
    int upload_set(const float* hx, const float* hy, const float* hm, uint32_t n) {
        int rc = x.put(hx, (size_t)n * fbm::NF);
        if (!rc) rc = y.put(hy, (size_t)n * fbm::NO);
        if (!rc) rc = mask.put(hm, (size_t)n * fbm::NO);
        return rc;
    }
};

namespace {

double mask_weight(const float* mask, const uint32_t* rows, uint32_t n) {
    double s = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t r = rows ? rows[i] : i;
        for (int o = 0; o < fbm::NO; ++o) s += std::fabs((double)mask[(size_t)r * fbm::NO + o]);
    }
    return s;
}

} // namespace

extern "C" {

int fbh_planner_new(uint64_t seed, fbh_planner** out) {
    FBH_API_LOCK();
    if (!out) return fail(FBH_ERR_ARG, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    fbh_planner* p = new fbh_planner();
    std::vector<float> w(fbm::NPARAM, 0.f);
    SplitMix r{seed};
    // He-uniform: U(-sqrt(6 / fan_in), +sqrt(6 / fan_in)) for each weight
    auto fill = [&](size_t off, int rows, int fan_in) {
        const double a = std::sqrt(6.0 / fan_in);
        for (size_t i = 0; i < (size_t)rows * fan_in; ++i) w[off + i] = (float)((2.0 * r.unit() - 1.0) * a);
    };
    const size_t oW1 = 0, oW2 = oW1 + fbm::H1 * fbm::NF + fbm::H1, oW3 = oW2 + fbm::H2 * fbm::H1 + fbm::H2;
    fill(oW1, fbm::H1, fbm::NF);
    fill(oW2, fbm::H2, fbm::H1);
    fill(oW3, fbm::NO, fbm::H2);
    std::vector<float> z(fbm::NPARAM, 0.f);
    for (int f = 0; f < fbm::NF; ++f) { p->hmu[f] = 0.f; p->hsd[f] = 1.f; }
    if ((rc = p->params.put(w.data(), fbm::NPARAM)) || (rc = p->m.put(z.data(), fbm::NPARAM)) ||
        (rc = p->v.put(z.data(), fbm::NPARAM)) || (rc = p->grad.ensure(fbm::NPARAM)) || (rc = p->put_norm()) ||
        (rc = check(hipStreamSynchronize(stream()), "sync"))) {
        delete p;
        return rc;
    }
    *out = p;
    return FBH_OK;
}

void fbh_planner_free(fbh_planner* p) {
    FBH_API_LOCK();
    delete p;
}

int fbh_planner_set_params(fbh_planner* p, const float* params, uint32_t n) {
    FBH_API_LOCK();
    if (!p || !params) return fail(FBH_ERR_ARG, "null argument");
    if (n != (uint32_t)fbm::NPARAM) return fail(FBH_ERR_SIZE, "parameter count is FBH_PLANNER_NPARAM");
    int rc = p->params.put(params, n);
    return rc ? rc : check(hipStreamSynchronize(stream()), "sync");
}

int fbh_planner_get_params(const fbh_planner* p, float* params, uint32_t n) {
    FBH_API_LOCK();
    if (!p || !params) return fail(FBH_ERR_ARG, "null argument");
    if (n != (uint32_t)fbm::NPARAM) return fail(FBH_ERR_SIZE, "parameter count is FBH_PLANNER_NPARAM");
    return p->params.get(params, n);
}

int fbh_planner_set_norm(fbh_planner* p, const float* mu, const float* sd) {
    FBH_API_LOCK();
    if (!p || !mu || !sd) return fail(FBH_ERR_ARG, "null argument");
    std::memcpy(p->hmu, mu, sizeof(p->hmu));
    std::memcpy(p->hsd, sd, sizeof(p->hsd));
    int rc = p->put_norm();
    return rc ? rc : check(hipStreamSynchronize(stream()), "sync");
}

int fbh_planner_get_norm(const fbh_planner* p, float* mu, float* sd) {
    FBH_API_LOCK();
    if (!p || !mu || !sd) return fail(FBH_ERR_ARG, "null argument");
    std::memcpy(mu, p->hmu, sizeof(p->hmu));
    std::memcpy(sd, p->hsd, sizeof(p->hsd));
    return FBH_OK;
}

int fbh_planner_forward(fbh_planner* p, const float* x, uint32_t n, float* out) {
    FBH_API_LOCK();
    if (!p || (n && (!x || !out))) return fail(FBH_ERR_ARG, "null argument");
    if (n == 0) return FBH_OK;
    int rc = p->x.put(x, (size_t)n * fbm::NF);
    if (!rc) rc = p->out.ensure((size_t)n * fbm::NO);
    if (!rc) rc = check(fbm::launch_fwd(p->params.p, p->x.p, nullptr, n, p->mu.p, p->sd.p, nullptr, nullptr, nullptr,
                                        p->out.p, stream()), "fwd");
    if (!rc) rc = p->out.get(out, (size_t)n * fbm::NO);
    return rc;
}

int fbh_planner_grad(fbh_planner* p, const float* x, const float* y, const float* mask, uint32_t n,
                     float* loss, float* grad) {
    FBH_API_LOCK();
    if (!p || !x || !y || !mask || !grad || n == 0) return fail(FBH_ERR_ARG, "null argument or no samples");
    const double mw = mask_weight(mask, nullptr, n);
    const float inv_m = mw > 0 ? (float)(1.0 / mw) : 0.f;
    int rc = p->upload_set(x, y, mask, n);
    if (!rc) rc = p->reserve(n);
    const uint32_t np = fbm::loss_parts(n);
    if (!rc) rc = p->loss.ensure(np);
    if (!rc) rc = p->fwd_delta(nullptr, n, inv_m, p->loss.p);
    if (!rc) rc = check(fbm::launch_grad(n, p->xn.p, p->h1.p, p->h2.p, p->d1.p, p->d2.p, p->d3.p, p->grad.p,
                                         stream()), "grad");
    std::vector<float> lp(np);
    if (!rc) rc = p->loss.get(lp.data(), np);
    if (!rc) rc = p->grad.get(grad, fbm::NPARAM);
    if (!rc && loss) {
        double s = 0;
        for (float v : lp) s += v;
        *loss = (float)s;
    }
    return rc;
}

int fbh_planner_step(fbh_planner* p, const float* x, const float* y, const float* mask, uint32_t n, float lr) {
    FBH_API_LOCK();
    if (!p || !x || !y || !mask || n == 0) return fail(FBH_ERR_ARG, "null argument or no samples");
    const double mw = mask_weight(mask, nullptr, n);
    const float inv_m = mw > 0 ? (float)(1.0 / mw) : 0.f;
    int rc = p->upload_set(x, y, mask, n);
    if (!rc) rc = p->reserve(n);
    if (!rc) rc = p->loss.ensure(fbm::loss_parts(n));
    if (!rc) rc = p->fwd_delta(nullptr, n, inv_m, p->loss.p);
    if (!rc) rc = check(fbm::launch_grad(n, p->xn.p, p->h1.p, p->h2.p, p->d1.p, p->d2.p, p->d3.p, p->grad.p,
                                         stream()), "grad");
    if (!rc) rc = p->adam(lr);
    if (!rc) rc = check(hipStreamSynchronize(stream()), "sync");
    return rc;
}

int fbh_planner_train(fbh_planner* p, const float* x, const float* y, const float* mask, uint32_t n,
                      uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float* epoch_loss) {
    FBH_API_LOCK();
    if (!p || !x || !y || !mask || n == 0 || batch == 0) return fail(FBH_ERR_ARG, "null argument or no samples");
    // normalisation and the in-range box from the training features
    float flo[fbm::NF], fhi[fbm::NF];
    for (int f = 0; f < fbm::NF; ++f) {
        double s = 0, s2 = 0, mn = x[f], mx = x[f];
        for (uint32_t i = 0; i < n; ++i) {
            const double v = x[(size_t)i * fbm::NF + f];
            s += v; s2 += v * v;
            mn = std::min(mn, v); mx = std::max(mx, v);
        }
        const double mean = s / n, var = std::max(0.0, s2 / n - mean * mean);
        p->hmu[f] = (float)mean;
        p->hsd[f] = var > 0 ? (float)std::sqrt(var) : 1.f;
        // 5% of the range as margin; a constant feature must match exactly
        const double mg = 0.05 * (mx - mn);
        flo[f] = (float)(mn - mg);
        fhi[f] = (float)(mx + mg);
    }
    const uint32_t b = std::min(batch, n);
    const uint32_t steps = (n + b - 1) / b;
    const uint32_t np = fbm::loss_parts(b);
    int rc = p->put_norm();
    if (!rc) rc = p->lo.put(flo, fbm::NF);
    if (!rc) rc = p->hi.put(fhi, fbm::NF);
    if (!rc) rc = p->upload_set(x, y, mask, n);
    if (!rc) rc = p->reserve(b);
    if (!rc) rc = p->idx.ensure(n);
    if (!rc) rc = p->loss.ensure((size_t)steps * np);
    if (rc) return rc;
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    SplitMix r{seed};
    std::vector<float> lp((size_t)steps * np);
    std::vector<double> mw(steps);
    for (uint32_t e = 0; e < epochs && !rc; ++e) {
        for (uint32_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[r.below(i + 1)]);   // Fisher-Yates
        // the previous epoch's work is done before perm is overwritten
        rc = p->idx.put(perm.data(), n);
        for (uint32_t st = 0; st < steps && !rc; ++st) {
            const uint32_t off = st * b, k = std::min(b, n - off);
            mw[st] = mask_weight(mask, perm.data() + off, k);
            const float inv_m = mw[st] > 0 ? (float)(1.0 / mw[st]) : 0.f;
            rc = p->fwd_delta(p->idx.p + off, k, inv_m, p->loss.p + (size_t)st * np);
            if (!rc) rc = check(fbm::launch_grad(k, p->xn.p, p->h1.p, p->h2.p, p->d1.p, p->d2.p, p->d3.p,
                                                 p->grad.p, stream()), "grad");
            if (!rc) rc = p->adam(lr);
        }
        if (!rc) rc = p->loss.get(lp.data(), (size_t)steps * np);
        if (!rc && epoch_loss) {
            // mean loss of the epoch, each batch weighted by its labels
            double num = 0, den = 0;
            for (uint32_t st = 0; st < steps; ++st) {
                const uint32_t k = std::min(b, n - st * b);
                double s = 0;
                for (uint32_t j = 0; j < fbm::loss_parts(k); ++j) s += lp[(size_t)st * np + j];
                num += s * mw[st];
                den += mw[st];
            }
            epoch_loss[e] = den > 0 ? (float)(num / den) : 0.f;
        }
    }
    if (!rc) p->trained = true;
    return rc;
}

int fbh_planner_choose(fbh_planner* p, const float* x, uint32_t n, const int* fallback, int* plans,
                       uint32_t* nfallback) {
    FBH_API_LOCK();
    if (!p || (n && (!x || !fallback || !plans))) return fail(FBH_ERR_ARG, "null argument");
    if (n == 0) { if (nfallback) *nfallback = 0; return FBH_OK; }
    if (!p->trained) {
        std::memcpy(plans, fallback, sizeof(int) * n);
        if (nfallback) *nfallback = n;
        return FBH_OK;
    }
    const uint32_t zero = 0;
    int rc = p->x.put(x, (size_t)n * fbm::NF);
    if (!rc) rc = p->fb.put(fallback, n);
    if (!rc) rc = p->plan.ensure(n);
    if (!rc) rc = p->out.ensure((size_t)n * fbm::NO);
    if (!rc) rc = p->nfall.put(&zero, 1);
    if (!rc) rc = check(fbm::launch_fwd(p->params.p, p->x.p, nullptr, n, p->mu.p, p->sd.p, nullptr, nullptr, nullptr,
                                        p->out.p, stream()), "fwd");
    if (!rc) rc = check(fbm::launch_choose(p->out.p, p->x.p, n, p->lo.p, p->hi.p, p->fb.p, p->plan.p, p->nfall.p,
                                           stream()), "choose");
    uint32_t nf = 0;
    if (!rc) rc = check(hipMemcpyAsync(&nf, p->nfall.p, 4, hipMemcpyDeviceToHost, stream()), "download");
    if (!rc) rc = p->plan.get(plans, n);
    if (!rc && nfallback) *nfallback = nf;
    return rc;
}

int fbh_index_choose_plans(fbh_index* ix, fbh_planner* p, uint32_t nq, const double* lo, const int* lop,
                           const double* hi, const int* rop, int* plans, uint32_t* nfallback) {
    FBH_API_LOCK();
    if (!ix || !p || (nq && (!lo || !lop || !hi || !rop || !plans))) return fail(FBH_ERR_ARG, "null argument");
    std::vector<float> f((size_t)nq * fbm::NF);
    std::vector<int> fb(nq);
    int rc = fbh_index_plan_features(ix, nq, lo, lop, hi, rop, f.data());
    for (uint32_t i = 0; i < nq && !rc; ++i) rc = fbh_index_estimate_cost(ix, lo[i], lop[i], hi[i], rop[i], nullptr,
                                                                         nullptr, &fb[i]);
    if (!rc) rc = fbh_planner_choose(p, f.data(), nq, fb.data(), plans, nfallback);
    return rc;
}

} // extern "C"
