// FP32 MLP cost model of the plan choice: 12 features -> 64 ReLU -> 64 ReLU
// -> 3 plan scores (predicted log2 of the plan's device time).  Forward,
// backward and Adam are hand-written kernels; the network is small enough
// that one thread carries one sample through a layer, with the weights in
// LDS, and one thread accumulates one parameter's gradient over the batch
// (no atomics, so training is deterministic).
//
// Parameter layout (flat): W1[64][12] b1[64] W2[64][64] b2[64] W3[3][64] b3[3].
#include "mlp_dev.h"

namespace fbm {

constexpr int TPB = 256;
constexpr int O_W1 = 0;
constexpr int O_B1 = O_W1 + H1 * NF;
constexpr int O_W2 = O_B1 + H1;
constexpr int O_B2 = O_W2 + H2 * H1;
constexpr int O_W3 = O_B2 + H2;
constexpr int O_B3 = O_W3 + NO * H2;
static_assert(O_B3 + NO == NPARAM, "parameter layout");

// Forward pass of samples idx[s] (or s when idx is null): normalised
// input xn (optional), hidden activations h1/h2 (optional, for backward)
// and outputs out[s][3].  The layer loops are unrolled so that the
// activations stay in registers.
__global__ __launch_bounds__(TPB) void k_mlp_fwd(const float* __restrict__ params, const float* __restrict__ x,
                                                 const uint32_t* __restrict__ idx, uint32_t n,
                                                 const float* __restrict__ mu, const float* __restrict__ sd,
                                                 float* __restrict__ xn, float* __restrict__ h1o,
                                                 float* __restrict__ h2o, float* __restrict__ out) {
    __shared__ float P[NPARAM];
    for (int i = threadIdx.x; i < NPARAM; i += TPB) P[i] = params[i];
    __syncthreads();
    const uint32_t s = blockIdx.x * TPB + threadIdx.x;
    if (s >= n) return;
    const uint32_t r = idx ? idx[s] : s;
    float in[NF];
    #pragma unroll
    for (int f = 0; f < NF; ++f) {
        const float d = sd[f] > 0.f ? sd[f] : 1.f;
        in[f] = (x[(size_t)r * NF + f] - mu[f]) / d;
        if (xn) xn[(size_t)s * NF + f] = in[f];
    }
    float a1[H1];
    #pragma unroll
    for (int j = 0; j < H1; ++j) {
        float acc = P[O_B1 + j];
        #pragma unroll
        for (int f = 0; f < NF; ++f) acc = fmaf(P[O_W1 + j * NF + f], in[f], acc);
        a1[j] = acc > 0.f ? acc : 0.f;
    }
    if (h1o) {
        #pragma unroll
        for (int j = 0; j < H1; ++j) h1o[(size_t)s * H1 + j] = a1[j];
    }
    float a2[H2];
    #pragma unroll
    for (int j = 0; j < H2; ++j) {
        float acc = P[O_B2 + j];
        #pragma unroll
        for (int k = 0; k < H1; ++k) acc = fmaf(P[O_W2 + j * H1 + k], a1[k], acc);
        a2[j] = acc > 0.f ? acc : 0.f;
    }
    if (h2o) {
        #pragma unroll
        for (int j = 0; j < H2; ++j) h2o[(size_t)s * H2 + j] = a2[j];
    }
    #pragma unroll
    for (int o = 0; o < NO; ++o) {
        float acc = P[O_B3 + o];
        #pragma unroll
        for (int k = 0; k < H2; ++k) acc = fmaf(P[O_W3 + o * H2 + k], a2[k], acc);
        out[(size_t)s * NO + o] = acc;
    }
}

// Per-sample deltas of the masked squared error (mask semantics in
// mlp_dev.h): d3 (3), d2 (64), d1 (64), and per-block partial losses.
__global__ __launch_bounds__(TPB) void k_mlp_delta(const float* __restrict__ params, const uint32_t* __restrict__ idx,
                                                   uint32_t n, const float* __restrict__ y,
                                                   const float* __restrict__ mask, float inv_m,
                                                   const float* __restrict__ h1, const float* __restrict__ h2,
                                                   const float* __restrict__ out, float* __restrict__ d1,
                                                   float* __restrict__ d2, float* __restrict__ d3,
                                                   float* __restrict__ loss_part) {
    __shared__ float P[NPARAM];
    __shared__ float red[TPB];
    for (int i = threadIdx.x; i < NPARAM; i += TPB) P[i] = params[i];
    __syncthreads();
    const uint32_t s = blockIdx.x * TPB + threadIdx.x;
    float l = 0.f;
    if (s < n) {
        const uint32_t r = idx ? idx[s] : s;
        float e3[NO];
        #pragma unroll
        for (int o = 0; o < NO; ++o) {
            const float m = mask[(size_t)r * NO + o];
            float e = out[(size_t)s * NO + o] - y[(size_t)r * NO + o];
            if (m < 0.f && e > 0.f) e = 0.f;          // lower bound met
            const float w = m < 0.f ? -m : m;
            l += w * e * e;
            e3[o] = 2.f * w * e * inv_m;
            d3[(size_t)s * NO + o] = e3[o];
        }
        float e2[H2];
        #pragma unroll
        for (int k = 0; k < H2; ++k) {
            float acc = 0.f;
            #pragma unroll
            for (int o = 0; o < NO; ++o) acc = fmaf(P[O_W3 + o * H2 + k], e3[o], acc);
            e2[k] = h2[(size_t)s * H2 + k] > 0.f ? acc : 0.f;
            d2[(size_t)s * H2 + k] = e2[k];
        }
        #pragma unroll
        for (int k = 0; k < H1; ++k) {
            float acc = 0.f;
            #pragma unroll
            for (int j = 0; j < H2; ++j) acc = fmaf(P[O_W2 + j * H1 + k], e2[j], acc);
            d1[(size_t)s * H1 + k] = h1[(size_t)s * H1 + k] > 0.f ? acc : 0.f;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] * inv_m;
}

// grad[p] = sum over the batch of delta[row(p)] * activation[col(p)]: one
// thread per parameter, samples in order (deterministic).
__global__ __launch_bounds__(TPB) void k_mlp_grad(uint32_t n, const float* __restrict__ xn,
                                                  const float* __restrict__ h1, const float* __restrict__ h2,
                                                  const float* __restrict__ d1, const float* __restrict__ d2,
                                                  const float* __restrict__ d3, float* __restrict__ grad) {
    const int p = blockIdx.x * TPB + threadIdx.x;
    if (p >= NPARAM) return;
    const float* dl;
    const float* act;
    int drow, dstride, acol, astride;
    if (p < O_B1) { drow = (p - O_W1) / NF; acol = (p - O_W1) % NF; dl = d1; dstride = H1; act = xn; astride = NF; }
    else if (p < O_W2) { drow = p - O_B1; acol = -1; dl = d1; dstride = H1; act = nullptr; astride = 0; }
    else if (p < O_B2) { drow = (p - O_W2) / H1; acol = (p - O_W2) % H1; dl = d2; dstride = H2; act = h1; astride = H1; }
    else if (p < O_W3) { drow = p - O_B2; acol = -1; dl = d2; dstride = H2; act = nullptr; astride = 0; }
    else if (p < O_B3) { drow = (p - O_W3) / H2; acol = (p - O_W3) % H2; dl = d3; dstride = NO; act = h2; astride = H2; }
    else { drow = p - O_B3; acol = -1; dl = d3; dstride = NO; act = nullptr; astride = 0; }
    float g = 0.f;
    #pragma unroll 8
    for (uint32_t s = 0; s < n; ++s) {
        const float d = dl[(size_t)s * dstride + drow];
        g = fmaf(d, acol < 0 ? 1.f : act[(size_t)s * astride + acol], g);
    }
    grad[p] = g;
}

// Adam (Kingma & Ba 2015) with bias correction; step t >= 1.
__global__ __launch_bounds__(TPB) void k_adam(float* __restrict__ params, const float* __restrict__ grad,
                                              float* __restrict__ m, float* __restrict__ v, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * TPB + threadIdx.x;
    if (p >= NPARAM) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// Inference: plan[s] = 1 + argmin of the three scores (LITERAL,
// COMPRESSED, SCAN; ties to the lower plan), or fallback[s] when a raw
// feature lies outside the training range [lo, hi].  nfall counts the
// fallbacks (one atomic per wavefront).
__global__ __launch_bounds__(TPB) void k_mlp_choose(const float* __restrict__ out, const float* __restrict__ x,
                                                    uint32_t n, const float* __restrict__ lo,
                                                    const float* __restrict__ hi,
                                                    const int32_t* __restrict__ fallback,
                                                    int32_t* __restrict__ plan, uint32_t* __restrict__ nfall) {
    const uint32_t s = blockIdx.x * TPB + threadIdx.x;
    bool fell = false;
    if (s < n) {
        int best = 0;
        float bv = out[(size_t)s * NO];
        #pragma unroll
        for (int o = 1; o < NO; ++o) {
            const float v = out[(size_t)s * NO + o];
            if (v < bv) { bv = v; best = o; }
        }
        if (lo && hi) {
            #pragma unroll
            for (int f = 0; f < NF; ++f) {
                const float v = x[(size_t)s * NF + f];
                fell |= !(v >= lo[f] && v <= hi[f]);
            }
        }
        plan[s] = fell ? fallback[s] : best + 1;
    }
    const uint64_t b = __ballot(fell);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(nfall, (uint32_t)__popcll(b));
}

static inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

hipError_t launch_fwd(const float* params, const float* x, const uint32_t* idx, uint32_t n, const float* mu,
                      const float* sd, float* xn, float* h1, float* h2, float* out, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_fwd, dim3(cdiv(n, TPB)), dim3(TPB), 0, s, params, x, idx, n, mu, sd, xn, h1, h2, out);
    return hipGetLastError();
}

hipError_t launch_delta(const float* params, const uint32_t* idx, uint32_t n, const float* y, const float* mask,
                        float inv_m, const float* h1, const float* h2, const float* out, float* d1, float* d2,
                        float* d3, float* loss_part, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_delta, dim3(cdiv(n, TPB)), dim3(TPB), 0, s, params, idx, n, y, mask, inv_m, h1, h2,
                       out, d1, d2, d3, loss_part);
    return hipGetLastError();
}

hipError_t launch_grad(uint32_t n, const float* xn, const float* h1, const float* h2, const float* d1,
                       const float* d2, const float* d3, float* grad, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(cdiv(NPARAM, TPB)), dim3(TPB), 0, s, n, xn, h1, h2, d1, d2, d3, grad);
    return hipGetLastError();
}

hipError_t launch_adam(float* params, const float* grad, float* m, float* v, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3(cdiv(NPARAM, TPB)), dim3(TPB), 0, s, params, grad, m, v, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_choose(const float* out, const float* x, uint32_t n, const float* lo, const float* hi,
                         const int32_t* fallback, int32_t* plan, uint32_t* nfall, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_choose, dim3(cdiv(n, TPB)), dim3(TPB), 0, s, out, x, n, lo, hi, fallback, plan, nfall);
    return hipGetLastError();
}

uint32_t loss_parts(uint32_t n) { return cdiv(n, TPB); }

} // namespace fbm
