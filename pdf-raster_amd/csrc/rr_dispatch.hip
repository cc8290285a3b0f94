// rr_dispatch.hip -- the learned dispatcher's MLP on the device: loading its
// parameters, a standalone forward kernel, and training (one workgroup,
// minibatch Adam on the masked squared error of predicted log2 span-pass
// times).  The in-kernel inference ahead of the fill is k_dispatch
// (rr_batch.hip), which shares mlp_forward (rr_mlp.h).
//
// The network is tiny (5341 parameters), so training keeps the parameters,
// gradients and one minibatch's activations in LDS and runs every step inside
// one persistent workgroup: no host round trip per step.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "rr_hip.h"
#include "rr_internal.h"
#include "rr_mlp.h"

using namespace rrdev;

static_assert(kMlpParams == RR_MLP_PARAMS, "parameter count");
static_assert(kFeatures == RR_MLP_FEATURES && kClasses == RR_MLP_CLASSES, "shape");

namespace {

constexpr int kTrainThreads = 256;
constexpr int kMaxBatch = 64;

__global__ void __launch_bounds__(256) k_mlp_forward(const float *__restrict__ params,
                                                     const float *__restrict__ feat, float *__restrict__ out,
                                                     long long n) {
    __shared__ float P[kMlpParams];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    const long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float o[kClasses];
    mlp_forward(P, feat + p * kFeatures, o);
    for (int c = 0; c < kClasses; ++c) out[p * kClasses + c] = o[c];
}

// splitmix64 finaliser as a counter-based generator
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float uniform01(uint64_t seed, uint64_t k) {
    return (static_cast<float>(mix64(seed ^ mix64(k)) >> 40) + 0.5f) * (1.0f / 16777216.0f);
}

// Box-Muller standard normal
__device__ __forceinline__ float normal(uint64_t seed, uint64_t k) {
    const float u1 = uniform01(seed, 2 * k), u2 = uniform01(seed, 2 * k + 1);
    return sqrtf(-2.f * logf(u1)) * cosf(6.28318530718f * u2);
}

struct TrainLds {
    float P[kMlpParams];
    float G[kMlpParams];
    float X[kMaxBatch][kFeatures];
    float T[kMaxBatch][kClasses];
    float H1[kMaxBatch][kHidden];
    float H2[kMaxBatch][kHidden];  // reused for the hidden-1 deltas
    float D2[kMaxBatch][kHidden];
    float O[kMaxBatch][kClasses];  // outputs, then output deltas
    float red[2 * kFeatures];
    float loss;
};

__global__ void __launch_bounds__(kTrainThreads) k_mlp_train(const float *__restrict__ feat,
                                                             const float *__restrict__ tgt, long long n,
                                                             int32_t steps, int32_t batch, float lr, uint64_t seed,
                                                             float *__restrict__ params, float *__restrict__ adam_m,
                                                             float *__restrict__ adam_v, float *__restrict__ loss_out) {
    extern __shared__ unsigned char lds_raw[];
    TrainLds &L = *reinterpret_cast<TrainLds *>(lds_raw);
    const int tid = threadIdx.x;
    // feature standardisation from the data (mean, standard deviation)
    float sum[kFeatures], sq[kFeatures];
    for (int i = 0; i < kFeatures; ++i) sum[i] = sq[i] = 0.f;
    for (long long s = tid; s < n; s += kTrainThreads)
        for (int i = 0; i < kFeatures; ++i) {
            const float v = feat[s * kFeatures + i];
            sum[i] += v;
            sq[i] += v * v;
        }
    if (tid < 2 * kFeatures) L.red[tid] = 0.f;
    __syncthreads();
    for (int i = 0; i < kFeatures; ++i) {
        atomicAdd(&L.red[i], sum[i]);
        atomicAdd(&L.red[kFeatures + i], sq[i]);
    }
    __syncthreads();
    const float inv_n = 1.f / static_cast<float>(n > 0 ? n : 1);
    for (int i = tid; i < kFeatures; i += kTrainThreads) {
        const float mu = L.red[i] * inv_n;
        const float var = fmaxf(L.red[kFeatures + i] * inv_n - mu * mu, 0.f);
        L.P[kMlpMu + i] = mu;
        L.P[kMlpSd + i] = fmaxf(sqrtf(var), 1e-3f);
    }
    // He-initialised weights; output biases at the mean valid target
    for (int j = kMlpW1 + tid; j < kMlpParams; j += kTrainThreads) {
        float w = 0.f;
        if (j < kMlpB1)
            w = normal(seed, j) * sqrtf(2.f / kFeatures);
        else if (j >= kMlpW2 && j < kMlpB2)
            w = normal(seed, j) * sqrtf(2.f / kHidden);
        else if (j >= kMlpW3 && j < kMlpB3)
            w = normal(seed, j) * sqrtf(1.f / kHidden);
        L.P[j] = w;
    }
    for (int j = tid; j < kMlpParams; j += kTrainThreads) adam_m[j] = adam_v[j] = 0.f;
    if (tid < kClasses) {
        float acc = 0.f, cnt = 0.f;
        for (long long s = 0; s < n; ++s) {
            const float t = tgt[s * kClasses + tid];
            if (!isnan(t)) {
                acc += t;
                cnt += 1.f;
            }
        }
        // a class with no timing in the data keeps a prediction of 2^64 us:
        // nothing supports choosing it (its gradients are all masked)
        L.P[kMlpB3 + tid] = cnt > 0.f ? acc / cnt : 64.f;
    }
    if (tid == 0) L.loss = 0.f;
    __syncthreads();

    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float b1t = 1.f, b2t = 1.f;
    for (int step = 0; step < steps; ++step) {
        // minibatch (with replacement), standardised inputs
        for (int e = tid; e < batch * (kFeatures + kClasses); e += kTrainThreads) {
            const int s = e / (kFeatures + kClasses), k = e % (kFeatures + kClasses);
            const long long idx =
                static_cast<long long>(mix64(seed ^ (static_cast<uint64_t>(step) << 20) ^ static_cast<uint64_t>(s)) %
                                       static_cast<uint64_t>(n));
            if (k < kFeatures)
                L.X[s][k] = (feat[idx * kFeatures + k] - L.P[kMlpMu + k]) / L.P[kMlpSd + k];
            else
                L.T[s][k - kFeatures] = tgt[idx * kClasses + (k - kFeatures)];
        }
        for (int j = tid; j < kMlpParams; j += kTrainThreads) L.G[j] = 0.f;
        __syncthreads();
        // forward
        for (int e = tid; e < batch * kHidden; e += kTrainThreads) {
            const int s = e / kHidden, u = e % kHidden;
            float acc = L.P[kMlpB1 + u];
            for (int i = 0; i < kFeatures; ++i) acc = fmaf(L.P[kMlpW1 + u * kFeatures + i], L.X[s][i], acc);
            L.H1[s][u] = fmaxf(acc, 0.f);
        }
        __syncthreads();
        for (int e = tid; e < batch * kHidden; e += kTrainThreads) {
            const int s = e / kHidden, u = e % kHidden;
            float acc = L.P[kMlpB2 + u];
            for (int k = 0; k < kHidden; ++k) acc = fmaf(L.P[kMlpW2 + u * kHidden + k], L.H1[s][k], acc);
            L.H2[s][u] = fmaxf(acc, 0.f);
        }
        __syncthreads();
        // outputs -> masked squared-error deltas
        float lsum = 0.f;
        for (int e = tid; e < batch * kClasses; e += kTrainThreads) {
            const int s = e / kClasses, c = e % kClasses;
            float acc = L.P[kMlpB3 + c];
            for (int u = 0; u < kHidden; ++u) acc = fmaf(L.P[kMlpW3 + c * kHidden + u], L.H2[s][u], acc);
            const float t = L.T[s][c];
            const float dlt = isnan(t) ? 0.f : acc - t;
            lsum += dlt * dlt;
            L.O[s][c] = 2.f * dlt / static_cast<float>(batch);
        }
        if (step >= steps - 100 && lsum != 0.f) atomicAdd(&L.loss, lsum / static_cast<float>(batch));
        __syncthreads();
        // output layer gradients; hidden-2 deltas
        for (int e = tid; e < kClasses * kHidden + kClasses; e += kTrainThreads) {
            if (e < kClasses * kHidden) {
                const int c = e / kHidden, u = e % kHidden;
                float acc = 0.f;
                for (int s = 0; s < batch; ++s) acc = fmaf(L.O[s][c], L.H2[s][u], acc);
                L.G[kMlpW3 + e] = acc;
            } else {
                const int c = e - kClasses * kHidden;
                float acc = 0.f;
                for (int s = 0; s < batch; ++s) acc += L.O[s][c];
                L.G[kMlpB3 + c] = acc;
            }
        }
        for (int e = tid; e < batch * kHidden; e += kTrainThreads) {
            const int s = e / kHidden, u = e % kHidden;
            float acc = 0.f;
            for (int c = 0; c < kClasses; ++c) acc = fmaf(L.O[s][c], L.P[kMlpW3 + c * kHidden + u], acc);
            L.D2[s][u] = L.H2[s][u] > 0.f ? acc : 0.f;
        }
        __syncthreads();
        // hidden-2 layer gradients; hidden-1 deltas (into H2)
        for (int e = tid; e < kHidden * kHidden; e += kTrainThreads) {
            const int u = e / kHidden, k = e % kHidden;
            float acc = 0.f;
            for (int s = 0; s < batch; ++s) acc = fmaf(L.D2[s][u], L.H1[s][k], acc);
            L.G[kMlpW2 + e] = acc;
        }
        for (int u = tid; u < kHidden; u += kTrainThreads) {
            float acc = 0.f;
            for (int s = 0; s < batch; ++s) acc += L.D2[s][u];
            L.G[kMlpB2 + u] = acc;
        }
        for (int e = tid; e < batch * kHidden; e += kTrainThreads) {
            const int s = e / kHidden, k = e % kHidden;
            float acc = 0.f;
            for (int u = 0; u < kHidden; ++u) acc = fmaf(L.D2[s][u], L.P[kMlpW2 + u * kHidden + k], acc);
            L.H2[s][k] = L.H1[s][k] > 0.f ? acc : 0.f;
        }
        __syncthreads();
        // hidden-1 layer gradients
        for (int e = tid; e < kHidden * kFeatures + kHidden; e += kTrainThreads) {
            if (e < kHidden * kFeatures) {
                const int u = e / kFeatures, i = e % kFeatures;
                float acc = 0.f;
                for (int s = 0; s < batch; ++s) acc = fmaf(L.H2[s][u], L.X[s][i], acc);
                L.G[kMlpW1 + e] = acc;
            } else {
                const int u = e - kHidden * kFeatures;
                float acc = 0.f;
                for (int s = 0; s < batch; ++s) acc += L.H2[s][u];
                L.G[kMlpB1 + u] = acc;
            }
        }
        __syncthreads();
        // Adam (the standardisation entries are not trained)
        b1t *= b1;
        b2t *= b2;
        const float c1 = 1.f / (1.f - b1t), c2 = 1.f / (1.f - b2t);
        for (int j = kMlpW1 + tid; j < kMlpParams; j += kTrainThreads) {
            const float g = L.G[j];
            const float m = b1 * adam_m[j] + (1.f - b1) * g;
            const float v = b2 * adam_v[j] + (1.f - b2) * g * g;
            adam_m[j] = m;
            adam_v[j] = v;
            L.P[j] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
        }
        __syncthreads();
    }
    for (int j = tid; j < kMlpParams; j += kTrainThreads) params[j] = L.P[j];
    if (tid == 0 && loss_out) *loss_out = L.loss / static_cast<float>(steps < 100 ? (steps > 0 ? steps : 1) : 100);
}

#define RR_TRY(ctx, expr)                                                                                          \
    do {                                                                                                           \
        hipError_t e_ = (expr);                                                                                    \
        if (e_ != hipSuccess) return rrint::fail(ctx, RR_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

extern "C" int32_t rr_hip_set_dispatcher(rr_hip_ctx *c, const float *params, int32_t n) {
    if (!c) return RR_E_INVARIANT;
    if (c->page_open) return rrint::fail(c, RR_E_INVARIANT, "set_dispatcher inside an open page");
    RR_TRY(c, hipSetDevice(c->device));
    ++c->ws_gen;  // captured graphs baked the previous dispatcher in
    if (!params) {
        c->mlp_loaded = false;
        return RR_OK;
    }
    if (n != kMlpParams) return rrint::fail(c, RR_E_INVARIANT, "set_dispatcher: expected RR_MLP_PARAMS floats");
    for (int i = 0; i < kFeatures; ++i)
        if (!(params[kMlpSd + i] > 0.f)) return rrint::fail(c, RR_E_INVARIANT, "set_dispatcher: feature scale <= 0");
    if (!c->mlp) RR_TRY(c, hipMalloc(&c->mlp, sizeof(float) * kMlpParams));
    RR_TRY(c, hipStreamSynchronize(c->stream));
    RR_TRY(c, hipMemcpy(c->mlp, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    c->mlp_loaded = true;
    return RR_OK;
}

extern "C" int32_t rr_hip_last_dispatch(rr_hip_ctx *c, float *out_features, uint8_t *out_routes, int64_t n_paths) {
    if (!c) return RR_E_INVARIANT;
    if (n_paths < 0 || (out_features && static_cast<size_t>(n_paths) * kFeatures > c->feat_cap) ||
        (out_routes && static_cast<size_t>(n_paths) > c->route_cls_cap))
        return rrint::fail(c, RR_E_INVARIANT, "last_dispatch: more paths than the last batched recording held");
    RR_TRY(c, hipSetDevice(c->device));
    RR_TRY(c, hipStreamSynchronize(c->stream));
    if (out_features && n_paths)
        RR_TRY(c, hipMemcpy(out_features, c->feat, sizeof(float) * kFeatures * n_paths, hipMemcpyDeviceToHost));
    if (out_routes && n_paths) RR_TRY(c, hipMemcpy(out_routes, c->route_cls, n_paths, hipMemcpyDeviceToHost));
    return RR_OK;
}

extern "C" int32_t rr_hip_mlp_forward(rr_hip_ctx *c, const float *features, float *out, int64_t n) {
    if (!c || n < 0 || (n > 0 && (!features || !out))) return RR_E_INVARIANT;
    if (!c->mlp_loaded) return rrint::fail(c, RR_E_INVARIANT, "mlp_forward: no dispatcher loaded");
    if (n == 0) return RR_OK;
    RR_TRY(c, hipSetDevice(c->device));
    hipLaunchKernelGGL(k_mlp_forward, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, c->stream, c->mlp,
                       features, out, static_cast<long long>(n));
    RR_TRY(c, hipGetLastError());
    RR_TRY(c, hipStreamSynchronize(c->stream));
    return RR_OK;
}

extern "C" int32_t rr_hip_train_dispatcher(rr_hip_ctx *c, const float *features, const float *targets, int64_t n,
                                           int32_t steps, int32_t batch, float lr, uint64_t seed, float *params_out,
                                           float *loss_out) {
    if (!c || !features || !targets || !params_out || n <= 0 || steps < 0 || batch <= 0 || batch > kMaxBatch ||
        !(lr > 0.f))
        return RR_E_INVARIANT;
    if (c->page_open) return rrint::fail(c, RR_E_INVARIANT, "train_dispatcher inside an open page");
    RR_TRY(c, hipSetDevice(c->device));
    float *buf = nullptr;  // params, Adam m, Adam v, loss
    RR_TRY(c, hipMalloc(&buf, sizeof(float) * (3 * kMlpParams + 1)));
    const size_t lds = sizeof(TrainLds);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mlp_train),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(kTrainThreads), lds, c->stream, features, targets,
                           static_cast<long long>(n), steps, batch, lr, seed, buf, buf + kMlpParams,
                           buf + 2 * kMlpParams, buf + 3 * kMlpParams);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(params_out, buf, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss_out) e = hipMemcpy(loss_out, buf + 3 * kMlpParams, sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess) return rrint::fail(c, RR_E_DEVICE, std::string("train_dispatcher: ") + hipGetErrorString(e));
    return RR_OK;
}
