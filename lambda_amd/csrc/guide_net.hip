// The guiding network's kernels -- inference, gradient, Adam, weight
// packing, tabulation -- and the host methods of GuideNet that launch them.
#include <cmath>
#include <vector>

#include "guide_net.h"

namespace lam {
namespace {

// Inference for n rows (min(*n_dev, n) when n_dev is given).  The block stages
// every layer's weights (W^T, 45 KB of bfloat16) in LDS once; its waves then
// take 64-row chunks, every (waves in the grid)-th one, with the activations
// in registers.  Without the staging each wave reloads all the weights
// through the vector L1 for every 64 rows, and that cache's access rate
// bounds the kernel.  ostride must be a multiple of 4.
__global__ void __launch_bounds__(kFwdThreads, 3) k_guide_forward(GuideDev G, const float *feat, int fstride, int n,
                                                               const int *n_dev, float *out, int ostride) {
    __shared__ __attribute__((aligned(16))) __bf16 sW[kGuideWeightsBf];
    for (int i = threadIdx.x * 8; i < kGuideWeightsBf; i += kFwdThreads * 8)
        *reinterpret_cast<bf16x8 *>(sW + i) = load8(G.w_nk + i);
    __syncthreads();
    // a count read on the device is capped at the rows the caller's buffers hold
    if (n_dev) n = min(*n_dev, n);
    GuideDev L = G;
    L.w_nk = sW;
    const long long step = (long long)gridDim.x * kFwdWaves * kFwdRowsPerWave;
    for (long long row0 = ((long long)blockIdx.x * kFwdWaves + (threadIdx.x >> 6)) * kFwdRowsPerWave; row0 < n;
         row0 += step) {
        // the weights are read from LDS in every chunk, not held in
        // registers across chunks
        asm volatile("" ::: "memory");
        mlp_rows64(L, feat + size_t(row0) * fstride, fstride, int(n - row0), out + size_t(row0) * ostride, ostride);
    }
}


// Gradient of sum_i -c_i log q(w_i) over samples [0, n), as per-block sums
// in `part` (one row of kGradRow4 float4 per block, layout below) that
// k_guide_reduce adds to grads (fp32, parameter layout) and *loss.
// Each block takes 64 rows: forward with stored activations, the loss
// gradient (four lanes per row), then backward through the five layers.  Weight
// gradients dW_l = H_{l-1}^T dZ_l sum over the block's rows on the MFMA
// (16 x 16 tiles shared among the block's waves).
// kChunks: blocks loop over several 64-row chunks (batches above
// kGradientBlocks chunks); otherwise each block takes one chunk.  The loop
// keeps the gradient registers live through the forward pass as well, which
// costs a small batch more than the rows it saves.
// Per-block sums instead of atomics: the blocks of a launch reach their
// 22,848 additions together, and same-address atomics from every block
// (and one loss atomic per wave, which the next weight loads waited
// behind) bounded the kernel.
template <bool kChunks>
__global__ void __launch_bounds__(256, 2) k_guide_gradient(GuideDev G, const float *samples, int n, float4 *part,
                                                        const float *c_cap) {
    __shared__ __attribute__((aligned(16))) __bf16 sX[kGuideRowsPerBlock * kLdX];
    __shared__ __attribute__((aligned(16))) __bf16 sH[4][kGuideRowsPerBlock * kLdH];   // H1..H4
    __shared__ __attribute__((aligned(16))) __bf16 sD[2][kGuideRowsPerBlock * kLdH];   // dZ ping-pong
    __shared__ __attribute__((aligned(16))) float sO[kGuideRowsPerBlock * kLobeOut];  // raw outputs (40 used)
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int r16 = l & 15, q = l >> 4;
    const int wr = w * kGuideRowsPerWave;  // this wave's first row in the block
    // The weight-gradient tiles, bias sums and loss stay in registers over
    // all of the block's 64-row chunks and are written once at the end.
    constexpr int kMaxTiles = kGuideIn / 16;   // tiles per wave of the widest layer
    f32x4 dw[kGuideLayers][kMaxTiles];
    float db[kGuideLayers];   // column l of dZ over this wave's 16 rows
    float lacc = 0;           // loss of this lane's rows
#pragma unroll
    for (int layer = 0; layer < kGuideLayers; layer++) {
#pragma unroll
        for (int i = 0; i < kMaxTiles; i++) dw[layer][i] = f32x4{0, 0, 0, 0};
        db[layer] = 0;
    }
    const int nchunks = (n + kGuideRowsPerBlock - 1) / kGuideRowsPerBlock;
    for (int chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const int row0 = chunk * kGuideRowsPerBlock;

    // forward
    encode_rows(samples, kGuideSample, row0 + wr, n, sX + wr * kLdX);
    __syncthreads();
    layer_rows16(G, 0, sX + wr * kLdX, kLdX, sH[0] + wr * kLdH, nullptr);
    __syncthreads();
    layer_rows16(G, 1, sH[0] + wr * kLdH, kLdH, sH[1] + wr * kLdH, nullptr);
    __syncthreads();
    layer_rows16(G, 2, sH[1] + wr * kLdH, kLdH, sH[2] + wr * kLdH, nullptr);
    __syncthreads();
    layer_rows16(G, 3, sH[2] + wr * kLdH, kLdH, sH[3] + wr * kLdH, nullptr);
    __syncthreads();
    layer_rows16(G, 4, sH[3] + wr * kLdH, kLdH, nullptr, sO + wr * kLobeOut);
    __syncthreads();

    // loss gradient: four lanes per row, two lobes each (the softmax and the
    // mixture density reduced over the four lanes; every thread of the
    // block takes part); dZ_4 in bf16, padding columns and rows >= n zero
    {
        const int r = threadIdx.x >> 2, part = threadIdx.x & 3;
        const bool valid = row0 + r < n;
        const float *sm = samples + size_t(valid ? row0 + r : 0) * kGuideSample;
        const V3 wd = v3(sm[6], sm[7], sm[8]);
        const float c = c_cap ? fminf(sm[9], *c_cap) : sm[9];
        const float *o0 = sO + r * kLobeOut + 10 * part, *o1 = o0 + 5;
        float mx = fmaxf(o0[4], o1[4]);
        mx = fmaxf(mx, __shfl_xor(mx, 1));
        mx = fmaxf(mx, __shfl_xor(mx, 2));
        const LobeEval e0 = guide_lobe_density(o0, wd), e1 = guide_lobe_density(o1, wd);
        float p0 = expf(o0[4] - mx), p1 = expf(o1[4] - mx);
        float sum = p0 + p1;
        sum += __shfl_xor(sum, 1);
        sum += __shfl_xor(sum, 2);
        p0 /= sum;
        p1 /= sum;
        float q = p0 * e0.comp + p1 * e1.comp;
        q += __shfl_xor(q, 1);
        q += __shfl_xor(q, 2);
        q = fmaxf(q, 1e-30f);
        float g[10];
        guide_lobe_grad(o0, wd, c, e0, p0, q, g);
        guide_lobe_grad(o1, wd, c, e1, p1, q, g + 5);
        __bf16 *dz = sD[0] + r * kLdH;
#pragma unroll
        for (int j = 0; j < 10; j++) dz[10 * part + j] = __bf16(valid ? g[j] : 0.0f);
#pragma unroll
        for (int j = 0; j < 6; j++) dz[kLobeOut + 6 * part + j] = __bf16(0.0f);
        if (valid && part == 0) lacc += -c * logf(q);
    }
    __syncthreads();

    int cur = 0;
    // unrolled: each layer's shapes are constants, so the products' loops
    // unroll and their weight loads are issued together
#pragma unroll
    for (int layer = kGuideLayers - 1; layer >= 0; layer--) {
        const int K = guide_layer_in(layer);
        const __bf16 *Hprev = layer == 0 ? sX : sH[layer - 1];
        const int ldp = layer == 0 ? kLdX : kLdH;
        const __bf16 *dZ = sD[cur];
        // dW_l[K x 64] = H_prev^T dZ over the block's 64 rows: tiles (kt, nt),
        // tile w + 4 i for this wave
        const int ktiles = K / 16;
#pragma unroll
        for (int i = 0; i < ktiles; i++) {
            const int t = w + kGuideWaves * i;
            const int kt = t >> 2, nt = t & 3;
            f32x4 acc = dw[layer][i];
#pragma unroll
            for (int rr = 0; rr < kGuideRowsPerBlock; rr += 32) {
                // A = H_prev^T: A[i][rows] = H_prev[rows][kt*16 + i]
                const bf16x8 a = load8_strided(Hprev + (rr + 8 * q) * ldp + kt * 16 + r16, ldp);
                // B = dZ: B[rows][j] = dZ[rows][nt*16 + j]
                const bf16x8 b = load8_strided(dZ + (rr + 8 * q) * kLdH + nt * 16 + r16, kLdH);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
            }
            dw[layer][i] = acc;
        }
        // db_l: column l of dZ over this wave's 16 rows
        {
            float sum = 0;
#pragma unroll
            for (int r = 0; r < kGuideRowsPerWave; r++) sum += float(dZ[(wr + r) * kLdH + l]);
            db[layer] += sum;
        }
        if (layer > 0) {
            // dH_prev[16 x 64] = dZ[16 x 64] W_l^T for this wave's rows; the
            // B operand W_l^T[n][k] is read n-major from W_l (k-major)
            f32x4 acc[4];
#pragma unroll
            for (int nt = 0; nt < 4; nt++) acc[nt] = f32x4{0, 0, 0, 0};
            gemm_rows16(dZ + wr * kLdH, kLdH, G.w_kn + guide_wbf_offset(layer), 64, 64, acc);
            // dZ_prev = dH_prev where H_prev > 0
            __bf16 *dn = sD[cur ^ 1];
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                const int c = nt * 16 + r16;
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const int row = wr + 4 * q + rr;
                    const float h = float(Hprev[row * ldp + c]);
                    dn[row * kLdH + c] = __bf16(h > 0 ? acc[nt][rr] : 0.0f);
                }
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    if constexpr (!kChunks) break;
    }  // chunks
    // this block's row of `part`: float4 slot j of thread x at j * 256 + x
    // (each store instruction writes 1 KB in one piece), then the bias
    // columns and the loss summed over the four waves through LDS (sO is
    // free once the last barrier has passed)
    float4 *row = part + size_t(blockIdx.x) * kGradRow4;
#pragma unroll
    for (int layer = 0; layer < kGuideLayers; layer++) {
        const int ktiles = guide_layer_in(layer) / 16;
#pragma unroll
        for (int i = 0; i < ktiles; i++) {
            const f32x4 v = dw[layer][i];
            row[(grad_slot(layer) + i) * 256 + threadIdx.x] = make_float4(v[0], v[1], v[2], v[3]);
        }
        sO[(layer * kGuideWaves + w) * 64 + l] = db[layer];
    }
    for (int off = 32; off > 0; off >>= 1) lacc += __shfl_xor(lacc, off);
    if (l == 0) sO[kGuideLayers * kGuideWaves * 64 + w] = lacc;
    __syncthreads();
    float *tail = reinterpret_cast<float *>(row + kGradSlots * 256);
    for (int e = threadIdx.x; e < kGuideLayers * 64; e += blockDim.x) {
        const float *s4 = sO + (e >> 6) * kGuideWaves * 64 + (e & 63);
        tail[e] = ((s4[0] + s4[64]) + s4[128]) + s4[192];
    }
    if (threadIdx.x == 0) {
        const float *s4 = sO + kGuideLayers * kGuideWaves * 64;
        tail[kGuideLayers * 64] = ((s4[0] + s4[1]) + s4[2]) + s4[3];
    }
}

// grads += the sum of the nb rows of `part` (and *loss += their loss).  A
// block takes 64 float4 columns and 64 rows (four groups of 16); its sums
// are added with one atomic per float per block.
__global__ void __launch_bounds__(256) k_guide_reduce(const float4 *part, int nb, float *grads, float *loss) {
    __shared__ float4 red[4][64];
    const int x = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int col = blockIdx.x * 64 + x;
    const int b0 = blockIdx.y * 64 + g * 16;
    float4 s = make_float4(0, 0, 0, 0);
    if (col < kGradRow4) {
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (b0 + k < nb) {
                const float4 v = part[size_t(b0 + k) * kGradRow4 + col];
                s.x += v.x;
                s.y += v.y;
                s.z += v.z;
                s.w += v.w;
            }
        }
    }
    red[g][x] = s;
    __syncthreads();
    if (g != 0 || col >= kGradRow4) return;
#pragma unroll
    for (int k = 1; k < 4; k++) {
        s.x += red[k][x].x;
        s.y += red[k][x].y;
        s.z += red[k][x].z;
        s.w += red[k][x].w;
    }
    const float v[4] = {s.x, s.y, s.z, s.w};
    if (col < kGradSlots * 256) {
        // weight-gradient slot j of thread t: tile (layer, i) of wave t / 64,
        // rows 4 q .. 4 q + 3 of column r16 (the MFMA's accumulator layout)
        const int j = col >> 8, t = col & 255;
        int layer = 0, i = j;
        if (j >= kGuideIn / 16) {
            layer = 1 + (j - kGuideIn / 16) / 4;
            i = (j - kGuideIn / 16) % 4;
        }
        const int w = t >> 6, q = (t & 63) >> 4, r16 = t & 15;
        const int tile = w + kGuideWaves * i, kt = tile >> 2, nt = tile & 3;
        float *gW = grads + guide_w_offset(layer);
#pragma unroll
        for (int rr = 0; rr < 4; rr++) atomicAdd(gW + (kt * 16 + 4 * q + rr) * 64 + nt * 16 + r16, v[rr]);
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int e = (col - kGradSlots * 256) * 4 + k;  // bias index layer * 64 + c, then the loss
            if (e < kGuideLayers * 64) atomicAdd(grads + guide_b_offset(e >> 6) + (e & 63), v[k]);
            else if (e == kGuideLayers * 64) atomicAdd(loss, v[k]);
        }
    }
}

// Adam step over all parameters with gradient scale `scale` (1 / batch
// weight); clears the gradients and rewrites the BF16 weight copies.
__global__ void k_guide_adam(float *params, float *grads, float *m, float *v, __bf16 *w_nk, __bf16 *w_kn,
                             float lr, float b1, float b2, float eps, float c1, float c2, float scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kGuideParams) return;
    const float g = grads[i] * scale;
    grads[i] = 0;
    const float mi = b1 * m[i] + (1 - b1) * g;
    const float vi = b2 * v[i] + (1 - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    float p = params[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    params[i] = p;
    // weight element? find its layer
    for (int layer = 0; layer < kGuideLayers; layer++) {
        const int wo = guide_w_offset(layer), K = guide_layer_in(layer);
        if (i >= wo && i < wo + K * 64) {
            const int k = (i - wo) / 64, c = (i - wo) % 64;
            w_kn[guide_wbf_offset(layer) + k * 64 + c] = __bf16(p);
            w_nk[guide_wbf_offset(layer) + c * K + k] = __bf16(p);
        }
    }
}

// Writes the BF16 copies from the master weights.
__global__ void k_guide_pack(const float *params, __bf16 *w_nk, __bf16 *w_kn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kGuideParams) return;
    const float p = params[i];
    for (int layer = 0; layer < kGuideLayers; layer++) {
        const int wo = guide_w_offset(layer), K = guide_layer_in(layer);
        if (i >= wo && i < wo + K * 64) {
            const int k = (i - wo) / 64, c = (i - wo) % 64;
            w_kn[guide_wbf_offset(layer) + k * 64 + c] = __bf16(p);
            w_nk[guide_wbf_offset(layer) + c * K + k] = __bf16(p);
        }
    }
}


// *sum += the weights of samples [0, n) (block sums, one atomic per block;
// *sum cleared before), then k_weight_cap turns it into factor x the mean.
__global__ void __launch_bounds__(256) k_weight_sum(const float *samples, int n, float *sum) {
    __shared__ float part[256];
    float s = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += samples[size_t(i) * kGuideSample + 9];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(sum, part[0]);
}

__global__ void k_weight_cap(float *cap, int n, float factor) { *cap = factor * *cap / float(n > 0 ? n : 1); }


// Features of table entries [0, n): each cell's centre with one of the six
// axis directions as the normal, at the table's wavelength.
__global__ void k_guide_table_features(GuideTableDev T, float3 box_lo, float3 box_hi, float *feat, long long n) {
    const V3 lo = v3(box_lo.x, box_lo.y, box_lo.z), hi = v3(box_hi.x, box_hi.y, box_hi.z);
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        V3 p, nn;
        guide_table_query(T, e, p, nn);
        const GuideFeatures g = guide_features(p, lo, hi, nn, T.lambda);
#pragma unroll
        for (int k = 0; k < kGuideScalars; k++) feat[e * kGuideScalars + k] = g.f[k];
    }
}

// Raw network outputs of rows [0, n) (stride kLobeOut) to the compact
// mixture, in place.
__global__ void k_guide_compact_rows(float *rows, long long n) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        guide_compact(rows + e * kLobeOut, rows + e * kLobeOut);
}


}  // namespace

void GuideNet::alloc() {
    params.alloc(kGuideParams);
    grads.alloc(kGuideParams);
    m.alloc(kGuideParams);
    v.alloc(kGuideParams);
    w_nk.alloc(kGuideWeightsBf);
    w_kn.alloc(kGuideWeightsBf);
    loss.alloc(1);
    cap.alloc(1);
    part.alloc(size_t(kGradientBlocks) * kGradRow4 * 4);
}

// He-uniform weights from a seeded PCG32, zero biases; the output layer
// starts small (near-uniform lobes) and its padding columns stay zero.
void GuideNet::init(uint64_t seed) {
    alloc();
    std::vector<float> p(kGuideParams, 0.0f);
    Pcg rng = pcg_seed(seed, 1);
    for (int layer = 0; layer < kGuideLayers; layer++) {
        const int K = guide_layer_in(layer);
        const float a = layer == kGuideLayers - 1 ? 0.1f * sqrtf(6.0f / K) : sqrtf(6.0f / K);
        for (int k = 0; k < K; k++)
            for (int c = 0; c < 64; c++) {
                float u = rng.next();
                p[size_t(guide_w_offset(layer) + k * 64 + c)] =
                    (layer == kGuideLayers - 1 && c >= kLobeOut) ? 0.0f : a * (2 * u - 1);
            }
    }
    set(p.data(), nullptr);
    reset_adam(nullptr);
}

void GuideNet::reset_adam(hipStream_t st) {
    HIP_OK(hipMemsetAsync(m.p, 0, kGuideParams * sizeof(float), st));
    HIP_OK(hipMemsetAsync(v.p, 0, kGuideParams * sizeof(float), st));
    HIP_OK(hipMemsetAsync(grads.p, 0, kGuideParams * sizeof(float), st));
    HIP_OK(hipStreamSynchronize(st));
    step = 0;
}

void GuideNet::set(const float *host, hipStream_t st) {
    HIP_OK(hipMemcpyAsync(params.p, host, kGuideParams * sizeof(float), hipMemcpyHostToDevice, st));
    k_guide_pack<<<(kGuideParams + 255) / 256, 256, 0, st>>>(params.p, w_nk.p, w_kn.p);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st));
}

void GuideNet::forward(const float *feat, int fstride, int n, float *out, int ostride, hipStream_t st) const {
    if (n <= 0) return;
    const int blocks = std::min((n + kFwdWaves * kFwdRowsPerWave - 1) / (kFwdWaves * kFwdRowsPerWave), kFwdMaxBlocks);
    k_guide_forward<<<blocks, kFwdThreads, 0, st>>>(dev(), feat, fstride, n, nullptr, out, ostride);
    HIP_OK(hipGetLastError());
}

// Inference for at most `bound` rows, the count read on the device.
void GuideNet::forward_dev(const float *feat, const int *n_dev, int bound, float *out, hipStream_t st) const {
    if (bound <= 0) return;
    const int blocks =
        std::min((bound + kFwdWaves * kFwdRowsPerWave - 1) / (kFwdWaves * kFwdRowsPerWave), kFwdMaxBlocks);
    k_guide_forward<<<blocks, kFwdThreads, 0, st>>>(dev(), feat, kGuideScalars, bound, n_dev, out, kLobeOut);
    HIP_OK(hipGetLastError());
}

// Accumulates the gradient of samples [0, n) into grads (weights capped
// at *c_cap when given).
void GuideNet::gradient(const float *samples, int n, hipStream_t st, const float *c_cap) {
    if (n <= 0) return;
    // one 64-row chunk per block, or kGradientBlocks blocks over several
    const int chunks = (n + kGuideRowsPerBlock - 1) / kGuideRowsPerBlock;
    const int blocks = std::min(chunks, kGradientBlocks);
    float4 *rows = reinterpret_cast<float4 *>(part.p);
    if (chunks > kGradientBlocks)
        k_guide_gradient<true><<<blocks, 256, 0, st>>>(dev(), samples, n, rows, c_cap);
    else
        k_guide_gradient<false><<<blocks, 256, 0, st>>>(dev(), samples, n, rows, c_cap);
    HIP_OK(hipGetLastError());
    k_guide_reduce<<<dim3((kGradRow4 + 63) / 64, (blocks + 63) / 64), 256, 0, st>>>(rows, blocks, grads.p, loss.p);
    HIP_OK(hipGetLastError());
}

void GuideNet::adam(float scale, hipStream_t st) {
    step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1 - powf(b1, float(step)), c2 = 1 - powf(b2, float(step));
    k_guide_adam<<<(kGuideParams + 255) / 256, 256, 0, st>>>(params.p, grads.p, m.p, v.p, w_nk.p, w_kn.p, lr, b1,
                                                             b2, 1e-8f, c1, c2, scale);
    HIP_OK(hipGetLastError());
}

// `epochs` passes of minibatch Adam over samples [0, n): steps of `batch`
// samples, weights capped at cap_factor x their mean over [0, n) when
// cap_factor > 0 (the cap computed once, before the first pass).
void GuideNet::train(const float *samples, int n, int batch, hipStream_t st, float cap_factor, int epochs) {
    if (n <= 0) return;
    const float *capp = nullptr;
    if (cap_factor > 0) {
        HIP_OK(hipMemsetAsync(cap.p, 0, sizeof(float), st));
        k_weight_sum<<<std::min((n + 255) / 256, 512), 256, 0, st>>>(samples, n, cap.p);
        HIP_OK(hipGetLastError());
        k_weight_cap<<<1, 1, 0, st>>>(cap.p, n, cap_factor);
        HIP_OK(hipGetLastError());
        capp = cap.p;
    }
    for (int e = 0; e < epochs; e++)
        for (int b0 = 0; b0 + batch <= n; b0 += batch) {
            gradient(samples + size_t(b0) * kGuideSample, batch, st, capp);
            adam(1.0f / float(batch), st);
        }
}

// The table's entries from the current weights: features of every entry
// into feat (entries x kGuideScalars floats), the network into rows
// (entries x kLobeOut), then the compact form in place.
void GuideNet::tabulate(const GuideTableDev &T, float3 box_lo, float3 box_hi, float *feat, float *rows,
                        long long entries, hipStream_t st) const {
    if (entries <= 0) return;
    const int blocks = int(std::min<long long>((entries + 255) / 256, 4096));
    k_guide_table_features<<<blocks, 256, 0, st>>>(T, box_lo, box_hi, feat, entries);
    HIP_OK(hipGetLastError());
    forward(feat, kGuideScalars, int(entries), rows, kLobeOut, st);
    k_guide_compact_rows<<<blocks, 256, 0, st>>>(rows, entries);
    HIP_OK(hipGetLastError());
}

}  // namespace lam
