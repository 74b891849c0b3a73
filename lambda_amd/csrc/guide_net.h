// The guiding network's host side: FP32 master parameters, Adam moments,
// gradient accumulator and the two bfloat16 weight copies in device memory,
// and the launches of its kernels (guide_net.hip).
#pragma once

#include "guide_mlp.h"
#include "runtime.h"

namespace lam {

struct GuideNet {
    DevBuf<float> params, grads, m, v, cap;
    DevBuf<__bf16> w_nk, w_kn;
    DevBuf<float> loss;       // summed loss of the last gradient or training call
    DevBuf<float> part;       // per-block gradient sums: kGradientBlocks rows of kGradRow4 float4
    int step = 0;
    float lr = 1e-3f;

    GuideDev dev() const { return GuideDev{params.p, w_nk.p, w_kn.p}; }

    void alloc();
    // He-uniform weights from a seeded PCG32, zero biases; the output layer
    // starts small (near-uniform lobes) and its padding columns stay zero.
    void init(uint64_t seed);
    void reset_adam(hipStream_t st);
    void set(const float *host, hipStream_t st);
    void forward(const float *feat, int fstride, int n, float *out, int ostride, hipStream_t st) const;
    // Inference for at most `bound` rows, the count read on the device.
    void forward_dev(const float *feat, const int *n_dev, int bound, float *out, hipStream_t st) const;
    // Accumulates the gradient of samples [0, n) into grads (weights capped
    // at *c_cap when given).
    void gradient(const float *samples, int n, hipStream_t st, const float *c_cap = nullptr);
    void adam(float scale, hipStream_t st);
    // `epochs` passes of minibatch Adam over samples [0, n): steps of `batch`
    // samples, weights capped at cap_factor x their mean over [0, n) when
    // cap_factor > 0.
    void train(const float *samples, int n, int batch, hipStream_t st, float cap_factor = kWeightCap, int epochs = 1);
    // The table's entries (compact mixtures) from the current weights; feat
    // is scratch of entries x kGuideScalars floats.
    void tabulate(const GuideTableDev &T, float3 box_lo, float3 box_hi, float *feat, float *rows, long long entries,
                  hipStream_t st) const;
};

}  // namespace lam
