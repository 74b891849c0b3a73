// The guiding network on the GPU: BF16 weights and activations on MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation), FP32 master weights, Adam.
// Included inside lambda_hip.hip's anonymous namespace, after its error and
// buffer helpers (Failure, HIP_OK, DevBuf).
//
// Every kernel works on 16-row tiles, one per wave: a layer is
// C[16 x 64] = A[16 x K] B[K x 64] as 4 column tiles x K/32 MFMA steps.
// Operand layout of the MFMA (checked on gfx950): lane l holds
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j = 0..7, and
// receives C[4 (l / 16) + r][l % 16], r = 0..3.
//
// Weights are kept twice in BF16: w_nk[l] = W_l^T (n-major, the B operand of
// the forward product) and w_kn[l] = W_l (k-major, the B operand of the
// backward product dH = dZ W^T); both are rewritten from the FP32 master
// copy after every Adam step.
#pragma once

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kGuideRowsPerWave = 16;
constexpr int kGuideWaves = 4;
constexpr int kGuideRowsPerBlock = kGuideRowsPerWave * kGuideWaves;  // 64
// LDS row strides (elements), padded by 8 to spread banks.
constexpr int kLdX = kGuideIn + 8;
constexpr int kLdH = kGuideHidden + 8;
// bf16 weight elements per layout: sum over layers of K_l * 64.
constexpr int kGuideWeightsBf = kGuideIn * 64 + 4 * 64 * 64;
static_assert(kGuideWeightsBf == LAMBDA_GUIDE_WEIGHTS_BF16, "bfloat16 weight count of the C ABI");

__host__ __device__ inline int guide_wbf_offset(int l) { return l == 0 ? 0 : kGuideIn * 64 + (l - 1) * 64 * 64; }

struct GuideDev {
    const float *params;     // fp32 master: W_l[K][64], b_l[64] per layer
    const __bf16 *w_nk;      // W_l^T [64][K_l] per layer
    const __bf16 *w_kn;      // W_l [K_l][64] per layer
};

// 8 consecutive bf16 from LDS or global memory (16-byte aligned).
__device__ __forceinline__ bf16x8 load8(const __bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// 8 bf16 with stride `stride` (elements) from LDS.
__device__ __forceinline__ bf16x8 load8_strided(const __bf16 *p, int stride) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = p[j * stride];
    return v;
}

// acc[nt] += A[16 x K] (LDS, row-major, stride lda) * B, B given n-major in
// Bt[n][k] (stride ldb), for the 4 column tiles nt.
__device__ __forceinline__ void gemm_rows16(const __bf16 *A, int lda, const __bf16 *Bt, int ldb, int K,
                                            f32x4 (&acc)[4]) {
    const int l = threadIdx.x & 63;
    const int r = l & 15, q = l >> 4;
    for (int kk = 0; kk < K; kk += 32) {
        const bf16x8 a = load8(A + r * lda + kk + 8 * q);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const bf16x8 b = load8(Bt + (nt * 16 + r) * ldb + kk + 8 * q);
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nt], 0, 0, 0);
        }
    }
}

// Encodes `rows` feature rows starting at row0 into X[rows][kLdX] (bf16);
// rows >= n are 0.  A lane encodes one (row, scalar) pair at a time: one
// sincos, 16 values, two 16-byte stores.
__device__ __forceinline__ void encode_rows(const float *feat, int fstride, int row0, int n, __bf16 *X,
                                            int rows = kGuideRowsPerWave) {
    const int l = threadIdx.x & 63;
    for (int e = l; e < rows * kGuideScalars; e += 64) {
        const int r = e / kGuideScalars, sc = e % kGuideScalars;
        float v[16];
        if (row0 + r < n) {
            guide_encode_scalar(feat[size_t(row0 + r) * fstride + sc], v);
        } else {
#pragma unroll
            for (int k = 0; k < 16; k++) v[k] = 0;
        }
        bf16x8 lo, hi;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            lo[k] = __bf16(v[k]);
            hi[k] = __bf16(v[8 + k]);
        }
        bf16x8 *dst = reinterpret_cast<bf16x8 *>(X + r * kLdX + sc * 16);
        dst[0] = lo;
        dst[1] = hi;
    }
}

// One layer for a wave's 16 rows: H_out = act(H_in W_l + b_l).  relu: hidden
// layers; otherwise the fp32 result goes to out (16 x 64, stride 64).
__device__ __forceinline__ void layer_rows16(const GuideDev &G, int layer, const __bf16 *Hin, int ldin,
                                             __bf16 *Hout, float *out) {
    const int l = threadIdx.x & 63;
    const int K = guide_layer_in(layer);
    f32x4 acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[nt] = f32x4{0, 0, 0, 0};
    gemm_rows16(Hin, ldin, G.w_nk + guide_wbf_offset(layer), K, K, acc);
    const float *bias = G.params + guide_b_offset(layer);
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const int c = nt * 16 + (l & 15);
        const float bc = bias[c];
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const int row = 4 * (l >> 4) + rr;
            const float v = acc[nt][rr] + bc;
            if (Hout) Hout[row * kLdH + c] = __bf16(fmaxf(v, 0.0f));
            else out[row * 64 + c] = v;
        }
    }
}

// Inference works on 64 rows per wave (4 waves per block).
constexpr int kFwdRowsPerWave = 64;

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x4 relu_bf16x4(f32x4 v, float4 b) {
    bf16x4 h;
    h[0] = __bf16(fmaxf(v[0] + b.x, 0.0f));
    h[1] = __bf16(fmaxf(v[1] + b.y, 0.0f));
    h[2] = __bf16(fmaxf(v[2] + b.z, 0.0f));
    h[3] = __bf16(fmaxf(v[3] + b.w, 0.0f));
    return h;
}

__device__ __forceinline__ bf16x8 cat_bf16x4(bf16x4 a, bf16x4 b) {
    return bf16x8{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
}

// The network for the 64 rows of one wave with the activations in
// registers (no LDS): every layer is computed transposed, H_next^T = W^T H^T,
// so that lane (r16 = l % 16, q = l / 16) receives H_next[rt*16 + r16][nt*16 +
// 4q + i], i = 0..3, for the 4 row tiles rt and 4 column tiles nt.  The next
// layer contracts over k in the order those registers already hold: its
// k-step m takes from lane (r16, q) the 8 values H[row][32m + 4q + 0..3] and
// H[row][32m + 16 + 4q + 0..3], and the weight fragment is read in the same
// order.  Layer 0 encodes its inputs in place: lane (r16, q) needs, at k-step
// m, bands 4 (q & 1) .. +3 of scalar 2m + (q >> 1) of each of its rows.
// Rows >= nvalid read zero features and are not written; out[row][0..39];
// row tiles of 16 with no valid row are skipped (nvalid is wave-uniform).
__device__ __forceinline__ void mlp_rows64(const GuideDev &G, const float *feat, int fstride, int nvalid, float *out,
                                           int ostride) {
    const int l = threadIdx.x & 63;
    const int r16 = l & 15, q = l >> 4;
    const int nrt = min(4, (nvalid + 15) >> 4);  // row tiles with valid rows (the same in every lane)
    f32x4 acc[4][4];
    bf16x8 hb[2][4];
    // layer 0: K = 96 in natural order, k = 32m + 8q + j; the inputs of one
    // k-step are encoded just before its products
    {
        const __bf16 *Wt = G.w_nk + guide_wbf_offset(0);  // [64 n][96]
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
#pragma unroll
            for (int rt = 0; rt < 4; rt++) acc[nt][rt] = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const int sc = 2 * m + (q >> 1);
            bf16x8 xb[4];
#pragma unroll
            for (int rt = 0; rt < 4; rt++) {
                if (rt >= nrt) break;
                const int row = rt * 16 + r16;
                const float x = row < nvalid ? feat[size_t(row) * fstride + sc] : 0.0f;
                float v[8];
                guide_encode_half(x, q & 1, v);
#pragma unroll
                for (int j = 0; j < 8; j++) xb[rt][j] = __bf16(v[j]);
            }
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                const bf16x8 a = load8(Wt + (nt * 16 + r16) * kGuideIn + 32 * m + 8 * q);
#pragma unroll
                for (int rt = 0; rt < 4; rt++)
                    if (rt < nrt) acc[nt][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, xb[rt], acc[nt][rt], 0, 0, 0);
            }
        }
    }
    // hidden layers 1..3 and the output layer 4 (its last column tile is
    // padding and is skipped)
#pragma unroll
    for (int layer = 1; layer < kGuideLayers; layer++) {
        const float *bprev = G.params + guide_b_offset(layer - 1);
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const float4 b0 = *reinterpret_cast<const float4 *>(bprev + 32 * m + 4 * q);
            const float4 b1 = *reinterpret_cast<const float4 *>(bprev + 32 * m + 16 + 4 * q);
#pragma unroll
            for (int rt = 0; rt < 4; rt++)
                hb[m][rt] = cat_bf16x4(relu_bf16x4(acc[2 * m][rt], b0), relu_bf16x4(acc[2 * m + 1][rt], b1));
        }
        const __bf16 *Wt = G.w_nk + guide_wbf_offset(layer);  // [64 n][64]
        const int ntiles = layer == kGuideLayers - 1 ? 3 : 4;
#pragma unroll
        for (int nt = 0; nt < ntiles; nt++) {
#pragma unroll
            for (int rt = 0; rt < 4; rt++) acc[nt][rt] = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const __bf16 *w = Wt + (nt * 16 + r16) * kGuideHidden + 32 * m + 4 * q;
                const bf16x8 a = cat_bf16x4(*reinterpret_cast<const bf16x4 *>(w),
                                            *reinterpret_cast<const bf16x4 *>(w + 16));
#pragma unroll
                for (int rt = 0; rt < 4; rt++)
                    if (rt < nrt) acc[nt][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hb[m][rt], acc[nt][rt], 0, 0, 0);
            }
        }
    }
    const float *bout = G.params + guide_b_offset(kGuideLayers - 1);
#pragma unroll
    for (int nt = 0; nt < 3; nt++) {
        const int c0 = nt * 16 + 4 * q;
        if (c0 >= kLobeOut) continue;
        const float4 b = *reinterpret_cast<const float4 *>(bout + c0);
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
            const int row = rt * 16 + r16;
            if (row < nvalid)
                *reinterpret_cast<float4 *>(out + size_t(row) * ostride + c0) =
                    make_float4(acc[nt][rt][0] + b.x, acc[nt][rt][1] + b.y, acc[nt][rt][2] + b.z, acc[nt][rt][3] + b.w);
        }
    }
}

// mlp_rows64 as a call of its own, for kernels whose other code is register
// bound (the persistent guided kernel): the caller's live registers are
// saved once around the call instead of the network's working set spilling
// into the caller's loops.
__device__ __attribute__((noinline)) void mlp_rows64_call(const GuideDev &G, const float *feat, int fstride,
                                                          int nvalid, float *out, int ostride) {
    mlp_rows64(G, feat, fstride, nvalid, out, ostride);
}

// Inference for n rows (min(*n_dev, n) when n_dev is given).  The block stages
// every layer's weights (W^T, 45 KB of bfloat16) in LDS once; its waves then
// take 64-row chunks, every (waves in the grid)-th one, with the activations
// in registers.  Without the staging each wave reloads all the weights
// through the vector L1 for every 64 rows, and that cache's access rate
// bounds the kernel.  ostride must be a multiple of 4.
constexpr int kFwdThreads = 256;
constexpr int kFwdWaves = kFwdThreads / 64;
constexpr int kFwdMaxBlocks = 768;  // three blocks per compute unit (3 waves/SIMD)

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

// Training sample: features[6], direction[3], weight c.
constexpr int kGuideSample = 10;

// Gradient of sum_i -c_i log q(w_i) over samples [0, n) into grads (fp32,
// parameter layout, accumulated with atomics) and the loss into *loss.
// Each block takes 64 rows: forward with stored activations, the loss
// gradient per row, then backward through the five layers.  Weight
// gradients dW_l = H_{l-1}^T dZ_l sum over the block's rows on the MFMA
// (16 x 16 tiles shared among the block's waves).
__global__ void __launch_bounds__(256) k_guide_gradient(GuideDev G, const float *samples, int n, float *grads,
                                                        float *loss, const float *c_cap) {
    __shared__ __attribute__((aligned(16))) __bf16 sX[kGuideRowsPerBlock * kLdX];
    __shared__ __attribute__((aligned(16))) __bf16 sH[4][kGuideRowsPerBlock * kLdH];   // H1..H4
    __shared__ __attribute__((aligned(16))) __bf16 sD[2][kGuideRowsPerBlock * kLdH];   // dZ ping-pong
    __shared__ float sO[kGuideRowsPerBlock * 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int r16 = l & 15, q = l >> 4;
    const int row0 = blockIdx.x * kGuideRowsPerBlock;
    const int wr = w * kGuideRowsPerWave;  // this wave's first row in the block

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
    layer_rows16(G, 4, sH[3] + wr * kLdH, kLdH, nullptr, sO + wr * 64);
    __syncthreads();

    // loss gradient per row (one lane per row), dZ_4 in bf16, padding zero
    float lsum = 0;
    if (threadIdx.x < kGuideRowsPerBlock) {
        const int r = threadIdx.x;
        float d[kLobeOut];
#pragma unroll
        for (int c = 0; c < kLobeOut; c++) d[c] = 0;
        if (row0 + r < n) {
            const float *sm = samples + size_t(row0 + r) * kGuideSample;
            const float c = c_cap ? fminf(sm[9], *c_cap) : sm[9];
            lsum = guide_loss_grad(sO + r * 64, v3(sm[6], sm[7], sm[8]), c, d);
        }
#pragma unroll
        for (int c = 0; c < 64; c++) sD[0][r * kLdH + c] = __bf16(c < kLobeOut ? d[c] : 0.0f);
    }
    // wave 0 holds all 64 rows' losses
    if (w == 0) {
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
        if (l == 0) atomicAdd(loss, lsum);
    }
    __syncthreads();

    int cur = 0;
    for (int layer = kGuideLayers - 1; layer >= 0; layer--) {
        const int K = guide_layer_in(layer);
        const __bf16 *Hprev = layer == 0 ? sX : sH[layer - 1];
        const int ldp = layer == 0 ? kLdX : kLdH;
        const __bf16 *dZ = sD[cur];
        float *gW = grads + guide_w_offset(layer);
        float *gb = grads + guide_b_offset(layer);
        // dW_l[K x 64] = H_prev^T dZ over the block's 64 rows: tiles (kt, nt)
        const int ktiles = K / 16;
        for (int t = w; t < ktiles * 4; t += kGuideWaves) {
            const int kt = t >> 2, nt = t & 3;
            f32x4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int rr = 0; rr < kGuideRowsPerBlock; rr += 32) {
                // A = H_prev^T: A[i][rows] = H_prev[rows][kt*16 + i]
                const bf16x8 a = load8_strided(Hprev + (rr + 8 * q) * ldp + kt * 16 + r16, ldp);
                // B = dZ: B[rows][j] = dZ[rows][nt*16 + j]
                const bf16x8 b = load8_strided(dZ + (rr + 8 * q) * kLdH + nt * 16 + r16, kLdH);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int k = kt * 16 + 4 * q + rr, c = nt * 16 + r16;
                atomicAdd(gW + k * 64 + c, acc[rr]);
            }
        }
        // db_l = column sums of dZ
        if (threadIdx.x < 64) {
            float s = 0;
            for (int r = 0; r < kGuideRowsPerBlock; r++) s += float(dZ[r * kLdH + threadIdx.x]);
            atomicAdd(gb + threadIdx.x, s);
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

// Training weights are single-path radiance estimates with heavy tails; a
// training call caps each at kWeightCap times their mean over the samples it
// trains on, so that one bright path does not steer a whole Adam step.  The
// cap shapes the learned proposal only: the renderer's estimator weighs
// every sample by its own density.
constexpr float kWeightCap = 20.0f;

// *cap = factor x the mean weight of samples [0, n) (one block).
__global__ void __launch_bounds__(1024) k_weight_cap(const float *samples, int n, float factor, float *cap) {
    __shared__ float part[1024];
    float s = 0;
    for (int i = threadIdx.x; i < n; i += 1024) s += samples[size_t(i) * kGuideSample + 9];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *cap = factor * part[0] / float(n > 0 ? n : 1);
}

// Host side of the network: parameters, Adam state and the BF16 copies.
struct GuideNet {
    DevBuf<float> params, grads, m, v, cap;
    DevBuf<__bf16> w_nk, w_kn;
    DevBuf<float> loss;       // summed loss of the last gradient or training call
    int step = 0;
    float lr = 1e-3f;

    GuideDev dev() const { return GuideDev{params.p, w_nk.p, w_kn.p}; }

    void alloc() {
        params.alloc(kGuideParams);
        grads.alloc(kGuideParams);
        m.alloc(kGuideParams);
        v.alloc(kGuideParams);
        w_nk.alloc(kGuideWeightsBf);
        w_kn.alloc(kGuideWeightsBf);
        loss.alloc(1);
        cap.alloc(1);
    }

    // He-uniform weights from a seeded PCG32, zero biases; the output layer
    // starts small (near-uniform lobes) and its padding columns stay zero.
    void init(uint64_t seed) {
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

    void reset_adam(hipStream_t st) {
        HIP_OK(hipMemsetAsync(m.p, 0, kGuideParams * sizeof(float), st));
        HIP_OK(hipMemsetAsync(v.p, 0, kGuideParams * sizeof(float), st));
        HIP_OK(hipMemsetAsync(grads.p, 0, kGuideParams * sizeof(float), st));
        HIP_OK(hipStreamSynchronize(st));
        step = 0;
    }

    void set(const float *host, hipStream_t st) {
        HIP_OK(hipMemcpyAsync(params.p, host, kGuideParams * sizeof(float), hipMemcpyHostToDevice, st));
        k_guide_pack<<<(kGuideParams + 255) / 256, 256, 0, st>>>(params.p, w_nk.p, w_kn.p);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(st));
    }

    void forward(const float *feat, int fstride, int n, float *out, int ostride, hipStream_t st) const {
        if (n <= 0) return;
        const int blocks = std::min((n + kFwdWaves * kFwdRowsPerWave - 1) / (kFwdWaves * kFwdRowsPerWave), kFwdMaxBlocks);
        k_guide_forward<<<blocks, kFwdThreads, 0, st>>>(dev(), feat, fstride, n, nullptr, out, ostride);
        HIP_OK(hipGetLastError());
    }

    // Inference for at most `bound` rows, the count read on the device.
    void forward_dev(const float *feat, const int *n_dev, int bound, float *out, hipStream_t st) const {
        if (bound <= 0) return;
        const int blocks =
            std::min((bound + kFwdWaves * kFwdRowsPerWave - 1) / (kFwdWaves * kFwdRowsPerWave), kFwdMaxBlocks);
        k_guide_forward<<<blocks, kFwdThreads, 0, st>>>(dev(), feat, kGuideScalars, bound, n_dev, out, kLobeOut);
        HIP_OK(hipGetLastError());
    }

    // Accumulates the gradient of samples [0, n) into grads (weights capped
    // at *c_cap when given).
    void gradient(const float *samples, int n, hipStream_t st, const float *c_cap = nullptr) {
        if (n <= 0) return;
        k_guide_gradient<<<(n + kGuideRowsPerBlock - 1) / kGuideRowsPerBlock, 256, 0, st>>>(dev(), samples, n,
                                                                                            grads.p, loss.p, c_cap);
        HIP_OK(hipGetLastError());
    }

    void adam(float scale, hipStream_t st) {
        step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1 - powf(b1, float(step)), c2 = 1 - powf(b2, float(step));
        k_guide_adam<<<(kGuideParams + 255) / 256, 256, 0, st>>>(params.p, grads.p, m.p, v.p, w_nk.p, w_kn.p, lr, b1,
                                                                 b2, 1e-8f, c1, c2, scale);
        HIP_OK(hipGetLastError());
    }

    // Minibatch Adam over samples [0, n): steps of `batch` samples, weights
    // capped at kWeightCap x their mean over [0, n).
    void train(const float *samples, int n, int batch, hipStream_t st) {
        if (n <= 0) return;
        k_weight_cap<<<1, 1024, 0, st>>>(samples, n, kWeightCap, cap.p);
        HIP_OK(hipGetLastError());
        for (int b0 = 0; b0 + batch <= n; b0 += batch) {
            gradient(samples + size_t(b0) * kGuideSample, batch, st, cap.p);
            adam(1.0f / float(batch), st);
        }
    }
};
