// rr_mlp.h -- forward pass of the dispatcher MLP (12 -> 64 -> 64 -> 5, ReLU,
// FP32), shared by the in-kernel dispatch pass and the standalone forward
// kernel.  P is the parameter array in the layout of rr_device.h (kMlp*).
#pragma once

#include <hip/hip_runtime.h>

#include "rr_device.h"

namespace rrdev {

__device__ __forceinline__ void mlp_forward(const float *P, const float *feat, float out[kClasses]) {
    float x[kFeatures];
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) x[i] = (feat[i] - P[kMlpMu + i]) / P[kMlpSd + i];
    float h1[kHidden];
#pragma unroll
    for (int u = 0; u < kHidden; ++u) {
        float acc = P[kMlpB1 + u];
#pragma unroll
        for (int i = 0; i < kFeatures; ++i) acc = fmaf(P[kMlpW1 + u * kFeatures + i], x[i], acc);
        h1[u] = fmaxf(acc, 0.f);
    }
#pragma unroll
    for (int c = 0; c < kClasses; ++c) out[c] = P[kMlpB3 + c];
    for (int u = 0; u < kHidden; ++u) {
        float acc = P[kMlpB2 + u];
#pragma unroll
        for (int k = 0; k < kHidden; ++k) acc = fmaf(P[kMlpW2 + u * kHidden + k], h1[k], acc);
        const float h2 = fmaxf(acc, 0.f);
#pragma unroll
        for (int c = 0; c < kClasses; ++c) out[c] = fmaf(P[kMlpW3 + c * kHidden + u], h2, out[c]);
    }
}

}  // namespace rrdev
