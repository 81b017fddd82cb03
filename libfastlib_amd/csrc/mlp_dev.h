// Shapes and launchers of the plan-choice MLP (12 -> 64 -> 64 -> 3, FP32).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbm {

constexpr int NF = 12;   // query features
constexpr int H1 = 64;
constexpr int H2 = 64;
constexpr int NO = 3;    // plan scores: LITERAL, COMPRESSED, SCAN
constexpr int NPARAM = H1 * NF + H1 + H2 * H1 + H2 + NO * H2 + NO;

hipError_t launch_fwd(const float* params, const float* x, const uint32_t* idx, uint32_t n, const float* mu,
                      const float* sd, float* xn, float* h1, float* h2, float* out, hipStream_t s);
// Masked squared error: mask > 0 an exact label of weight mask, mask < 0 a
// lower bound of weight -mask (penalised only when the score is below it),
// 0 no label.  inv_m = 1 / sum |mask| over the batch.
hipError_t launch_delta(const float* params, const uint32_t* idx, uint32_t n, const float* y, const float* mask,
                        float inv_m, const float* h1, const float* h2, const float* out, float* d1, float* d2,
                        float* d3, float* loss_part, hipStream_t s);
hipError_t launch_grad(uint32_t n, const float* xn, const float* h1, const float* h2, const float* d1,
                       const float* d2, const float* d3, float* grad, hipStream_t s);
hipError_t launch_adam(float* params, const float* grad, float* m, float* v, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t s);
// plan[s] = 1 + argmin of the scores, or fallback[s] when a feature of
// sample s lies outside [lo, hi] (lo/hi null: never).
hipError_t launch_choose(const float* out, const float* x, uint32_t n, const float* lo, const float* hi,
                         const int32_t* fallback, int32_t* plan, uint32_t* nfall, hipStream_t s);
uint32_t loss_parts(uint32_t n);

} // namespace fbm
