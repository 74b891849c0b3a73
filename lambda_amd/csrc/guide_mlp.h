// The guiding network on the GPU: BF16 weights and activations on MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation), FP32 master weights, Adam.
// Device helpers and constants; the kernels and the host-side network
// (GuideNet) are in guide_net.hip / guide_net.h.
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

#include "guide.h"
#include "lambda_hip.h"

namespace lam {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kGuideRowsPerWave = 16;
constexpr int kGuideWaves = 4;
constexpr int kGuideRowsPerBlock = kGuideRowsPerWave * kGuideWaves;  // 64
// Gradient kernel grid: at most this many blocks (two per compute unit at
// 77 KB of LDS each on 256 compute units), each over several 64-row chunks.
constexpr int kGradientBlocks = 512;
// A gradient block's row of sums: the weight-gradient tiles as its 256
// threads hold them (kGradSlots float4 each: six tiles of the input layer,
// four of each other layer), the 5 x 64 bias sums, the loss, padding.
constexpr int kGradSlots = kGuideIn / 16 + 4 * (kGuideLayers - 1);
constexpr int kGradRow4 = kGradSlots * 256 + (kGuideLayers * 64 + 1 + 3) / 4;
// slot of tile i of layer l
__host__ __device__ constexpr int grad_slot(int l) { return l == 0 ? 0 : kGuideIn / 16 + 4 * (l - 1); }
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
#pragma unroll
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
// layers; otherwise the fp32 result's first kLobeOut columns go to out
// (16 x kLobeOut, stride kLobeOut).
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
            else if (c < kLobeOut) out[row * kLobeOut + c] = v;
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

constexpr int kFwdThreads = 256;
constexpr int kFwdWaves = kFwdThreads / 64;
constexpr int kFwdMaxBlocks = 768;  // three blocks per compute unit (3 waves/SIMD)

// Training sample: features[6], direction[3], weight c.
constexpr int kGuideSample = 10;

// Training weights are single-path radiance estimates with heavy tails; a
// training call caps each at kWeightCap times their mean over the samples it
// trains on, so that one bright path does not steer a whole Adam step.  The
// cap shapes the learned proposal only: the renderer's estimator weighs
// every sample by its own density.
constexpr float kWeightCap = 20.0f;

}  // namespace lam
