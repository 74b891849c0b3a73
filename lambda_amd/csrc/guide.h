// Path-guiding distribution: a mixture of von Mises-Fisher lobes on the
// sphere whose parameters come from a small MLP of (position, direction,
// wavelength).  Host and device code (the CPU checks include it).
//
// Network: input = 6 scalars, each frequency-encoded with 8 bands as
// sin(2^b pi x), cos(2^b pi x) (96 values); 4 hidden layers of 64 ReLU units;
// output 64 values of which the first 40 are 8 lobes x {m.xyz, kappa_raw,
// weight_logit}; the other 24 are padding kept at zero.
//
// Lobe k: mean mu_k = m_k / |m_k|, concentration kappa_k = exp(kappa_raw)
// clamped to [kKappaMin, kKappaMax] (log scale: the concentrations a narrow
// opening needs, hundreds, are a few Adam steps from the initial 1), weight
// pi_k = softmax(weight_logit)_k.
// Density of a lobe: vMF(w) = kappa / (2 pi (1 - exp(-2 kappa))) *
// exp(kappa (mu.w - 1)), which tends to 1 / (4 pi) as kappa -> 0.
#pragma once

#include "dmath.h"

namespace lam {

constexpr int kGuideScalars = 6;
constexpr int kGuideBands = 8;
constexpr int kGuideIn = kGuideScalars * kGuideBands * 2;  // 96
constexpr int kGuideHidden = 64;
constexpr int kGuideLayers = 5;                            // 4 hidden + output
constexpr int kGuideOut = 64;                              // padded output width
constexpr int kLobes = 8;
constexpr int kLobeOut = 5 * kLobes;                       // 40 used outputs
constexpr float kKappaMin = 1e-3f;
constexpr float kKappaMax = 2000.0f;

// Parameter layout (fp32): for each layer l, W_l[K_l][64] row-major then
// b_l[64]; K_0 = 96, K_1..4 = 64.
LAM_FN int guide_layer_in(int l) { return l == 0 ? kGuideIn : kGuideHidden; }
LAM_FN int guide_w_offset(int l) { return l == 0 ? 0 : kGuideIn * 64 + 64 + (l - 1) * (64 * 64 + 64); }
LAM_FN int guide_b_offset(int l) { return guide_w_offset(l) + guide_layer_in(l) * 64; }
constexpr int kGuideParams = kGuideIn * 64 + 64 + 4 * (64 * 64 + 64);  // 22848

// Raw features of one query, each mapped to [0, 1]: position in the scene
// box, a direction as (theta / pi, (phi + pi) / 2pi), wavelength over
// 360-830 nm.  The renderer guides Lambertian vertices only, where the
// integrand L_i cos / pi does not depend on the outgoing direction, and
// passes the surface normal (on the side of the incoming ray) as the
// direction: the cosine factor of the learned product depends on it.
struct GuideFeatures {
    float f[kGuideScalars];
};

LAM_FN GuideFeatures guide_features(V3 p, V3 lo, V3 hi, V3 dir, float lambda) {
    GuideFeatures g;
    g.f[0] = (p.x - lo.x) / fmaxf(hi.x - lo.x, 1e-6f);
    g.f[1] = (p.y - lo.y) / fmaxf(hi.y - lo.y, 1e-6f);
    g.f[2] = (p.z - lo.z) / fmaxf(hi.z - lo.z, 1e-6f);
    g.f[3] = acosf(fminf(fmaxf(dir.y, -1.0f), 1.0f)) / kPi;
    g.f[4] = (atan2f(dir.z, dir.x) + kPi) / (2 * kPi);
    g.f[5] = (lambda - 360.0f) / 470.0f;
    return g;
}

// Encoded input value i (0..95): scalar i / 16, band (i % 16) / 2, sin for
// even i, cos for odd i.
LAM_FN float guide_encode(const GuideFeatures &g, int i) {
    const int s = i >> 4, b = (i & 15) >> 1;
    const float x = g.f[s] * float(1 << b) * kPi;
    return (i & 1) ? cosf(x) : sinf(x);
}

// All 16 encoded values of scalar s (bands 0..7, sin then cos) from one
// sincos and the double-angle recurrences sin 2x = 2 sin x cos x,
// cos 2x = 1 - 2 sin^2 x; the recurrence error (about 1e-5 after 7
// doublings) is far below the bfloat16 resolution the network stores.
LAM_FN void guide_encode_scalar(float x, float (&v)[16]) {
    float sn, cs;
    sincosf(x * kPi, &sn, &cs);
#pragma unroll
    for (int b = 0; b < kGuideBands; b++) {
        v[2 * b] = sn;
        v[2 * b + 1] = cs;
        const float s2 = 2 * sn * cs;
        cs = 1 - 2 * sn * sn;
        sn = s2;
    }
}

// Half of the encoding of one scalar: bands 4h .. 4h + 3 (h = 0 or 1), sin
// then cos per band, from one sincos of 2^(4h) pi x and three double-angle
// steps (what one lane of the inference kernel feeds the matrix core).
LAM_FN void guide_encode_half(float x, int h, float (&v)[8]) {
    float sn, cs;
    sincosf(x * (h ? 16.0f * kPi : kPi), &sn, &cs);
#pragma unroll
    for (int b = 0; b < 4; b++) {
        v[2 * b] = sn;
        v[2 * b + 1] = cs;
        const float s2 = 2 * sn * cs;
        cs = 1 - 2 * sn * sn;
        sn = s2;
    }
}

struct Lobe {
    V3 mu;
    float kappa, norm, weight;  // norm = kappa / (2 pi (1 - exp(-2 kappa)))
};

LAM_FN float vmf_norm(float kappa) { return kappa / (2 * kPi * -expm1f(-2 * kappa)); }

// Lobes from the 40 raw outputs.
LAM_FN void guide_lobes(const float *o, Lobe (&L)[kLobes]) {
    float mx = -INFINITY;
    for (int k = 0; k < kLobes; k++) mx = fmaxf(mx, o[5 * k + 4]);
    float sum = 0;
    for (int k = 0; k < kLobes; k++) {
        const float *q = o + 5 * k;
        V3 m = v3(q[0], q[1], q[2]);
        float l = fmaxf(len(m), 1e-6f);
        L[k].mu = m * (1 / l);
        L[k].kappa = fminf(fmaxf(expf(q[3]), kKappaMin), kKappaMax);
        L[k].norm = vmf_norm(L[k].kappa);
        L[k].weight = expf(q[4] - mx);
        sum += L[k].weight;
    }
    for (int k = 0; k < kLobes; k++) L[k].weight /= sum;
}

LAM_FN float guide_pdf(const Lobe (&L)[kLobes], V3 w) {
    float p = 0;
    for (int k = 0; k < kLobes; k++) p += L[k].weight * L[k].norm * expf(L[k].kappa * (dot(L[k].mu, w) - 1));
    return p;
}

// Orthonormal frame around a unit vector (Duff et al., "Building an
// orthonormal basis, revisited", JCGT 2017).
LAM_FN void onb(V3 n, V3 &t, V3 &b) {
    float s = copysignf(1.0f, n.z);
    float a = -1 / (s + n.z);
    float c = n.x * n.y * a;
    t = v3(1 + s * n.x * n.x * a, s * c, -s * n.x);
    b = v3(c, s + n.y * n.y * a, -n.y);
}

// Sample of one vMF lobe from two uniforms (W. Jakob, "Numerically stable
// sampling of the von Mises Fisher distribution on S^2", 2012):
// cos theta = 1 + log(1 + u1 (exp(-2 kappa) - 1)) / kappa.
LAM_FN V3 vmf_sample(const Lobe &L, float u1, float u2) {
    float w = 1 + log1pf(u1 * expm1f(-2 * L.kappa)) / L.kappa;
    w = fminf(fmaxf(w, -1.0f), 1.0f);
    float v = sqrtf(fmaxf(0.0f, 1 - w * w));
    float phi = 2 * kPi * u2;
    V3 t, b;
    onb(L.mu, t, b);
    return normalize(t * (v * cosf(phi)) + b * (v * sinf(phi)) + L.mu * w);
}

// Sample of the mixture: u0 picks the lobe, u1 and u2 sample it.
LAM_FN V3 guide_sample(const Lobe (&L)[kLobes], float u0, float u1, float u2) {
    int k = 0;
    float c = L[0].weight;
    while (k < kLobes - 1 && u0 >= c) c += L[++k].weight;
    return vmf_sample(L[k], u1, u2);
}

// Compact form of the mixture at one vertex, 5 floats per lobe (40 in all,
// the size of the raw outputs): kappa mu (xyz), log(weight norm) - kappa, and
// the cumulative weight through the lobe, so that
//   q(w) = sum_k exp(c_k + (kappa mu)_k . w)
// takes one exp per lobe and sampling needs no other state.  c may alias o.
LAM_FN void guide_compact(const float *o, float *c) {
    Lobe L[kLobes];
    guide_lobes(o, L);
    float cum = 0;
#pragma unroll
    for (int k = 0; k < kLobes; k++) {
        cum += L[k].weight;
        c[5 * k + 0] = L[k].kappa * L[k].mu.x;
        c[5 * k + 1] = L[k].kappa * L[k].mu.y;
        c[5 * k + 2] = L[k].kappa * L[k].mu.z;
        c[5 * k + 3] = logf(L[k].weight * L[k].norm) - L[k].kappa;
        c[5 * k + 4] = cum;
    }
}

// The compact mixture's 40 floats in ten 16-byte loads (c is 16-byte
// aligned: rows of kLobeOut floats from an aligned base).
LAM_FN void guide_load_c(const float *c, float (&v)[kLobeOut]) {
    const float4 *c4 = reinterpret_cast<const float4 *>(c);
#pragma unroll
    for (int i = 0; i < kLobeOut / 4; i++) {
        const float4 q = c4[i];
        v[4 * i] = q.x;
        v[4 * i + 1] = q.y;
        v[4 * i + 2] = q.z;
        v[4 * i + 3] = q.w;
    }
}

LAM_FN float guide_pdf_c(const float *c, V3 w) {
    float v[kLobeOut];
    guide_load_c(c, v);
    float p = 0;
#pragma unroll
    for (int k = 0; k < kLobes; k++) p += expf(v[5 * k + 3] + (v[5 * k] * w.x + v[5 * k + 1] * w.y + v[5 * k + 2] * w.z));
    return p;
}

LAM_FN V3 guide_sample_c(const float *c, float u0, float u1, float u2) {
    // lobe k: the number of cumulative weights at or below u0 (they rise
    // with k), counted without a data-dependent loop
    int k = 0;
#pragma unroll
    for (int j = 0; j < kLobes - 1; j++) k += u0 >= c[5 * j + 4] ? 1 : 0;
    Lobe L;
    const V3 km = v3(c[5 * k], c[5 * k + 1], c[5 * k + 2]);
    L.kappa = len(km);
    L.mu = km * (1 / L.kappa);
    return vmf_sample(L, u1, u2);
}

// Tabulated guide: the network evaluated once per training step at the
// centres of a grid of cubic cells over the scene box, for each of the six
// axis directions as the normal, and stored in compact form (kLobeOut floats
// per entry, entries [z][y][x][normal]).  A vertex takes the entry of the
// cell that holds it and of the axis direction nearest its normal; the
// mixture it samples and the densities its MIS weights use are both that
// entry's, so the estimator stays unbiased whatever the table holds.
constexpr int kTableNormals = 6;

struct GuideTableDev {
    const float *rows;
    float lo[3];
    float inv_cell;  // cells per unit length
    int n[3];
    float lambda;    // wavelength the entries were evaluated at
};

// Normal bin: 2 * axis + (component < 0) for the axis of the largest
// component (ties to the lower axis).
LAM_FN int guide_table_normal_bin(V3 n) {
    const float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    if (ax >= ay && ax >= az) return n.x < 0 ? 1 : 0;
    if (ay >= az) return n.y < 0 ? 3 : 2;
    return n.z < 0 ? 5 : 4;
}

LAM_FN V3 guide_table_normal(int bin) {
    const float s = (bin & 1) ? -1.0f : 1.0f;
    return bin < 2 ? v3(s, 0, 0) : bin < 4 ? v3(0, s, 0) : v3(0, 0, s);
}

LAM_FN int guide_table_cell(float x, float lo, float inv_cell, int n) {
    const float f = floorf((x - lo) * inv_cell);
    return !(f > 0) ? 0 : f >= float(n - 1) ? n - 1 : int(f);  // NaN -> 0
}

LAM_FN long long guide_table_entry(const GuideTableDev &T, V3 p, V3 n) {
    const int ix = guide_table_cell(p.x, T.lo[0], T.inv_cell, T.n[0]);
    const int iy = guide_table_cell(p.y, T.lo[1], T.inv_cell, T.n[1]);
    const int iz = guide_table_cell(p.z, T.lo[2], T.inv_cell, T.n[2]);
    return ((((long long)iz * T.n[1] + iy) * T.n[0] + ix) * kTableNormals + guide_table_normal_bin(n));
}

// Centre and normal of entry e.
LAM_FN void guide_table_query(const GuideTableDev &T, long long e, V3 &p, V3 &n) {
    const int bin = int(e % kTableNormals);
    long long c = e / kTableNormals;
    const int ix = int(c % T.n[0]);
    c /= T.n[0];
    const int iy = int(c % T.n[1]);
    const int iz = int(c / T.n[1]);
    const float cell = 1 / T.inv_cell;
    p = v3(T.lo[0] + (ix + 0.5f) * cell, T.lo[1] + (iy + 0.5f) * cell, T.lo[2] + (iz + 0.5f) * cell);
    n = guide_table_normal(bin);
}

// Loss of one training sample, -c log q(w), and its gradient, split by lobe
// so that the lobes of one row can be spread over lanes: guide_lobe_density
// gives lobe k's vMF density at w (with its mean and concentration), the
// caller forms the softmax weights pi_k and q = sum_k pi_k vMF_k(w), and
// guide_lobe_grad gives the gradient with respect to lobe k's 5 raw outputs.
struct LobeEval {
    V3 mu;
    float kappa, comp;  // concentration, vMF_k(w)
};

LAM_FN LobeEval guide_lobe_density(const float *r, V3 w) {
    LobeEval e;
    V3 m = v3(r[0], r[1], r[2]);
    float l = fmaxf(len(m), 1e-6f);
    e.mu = m * (1 / l);
    e.kappa = fminf(fmaxf(expf(r[3]), kKappaMin), kKappaMax);
    e.comp = vmf_norm(e.kappa) * expf(e.kappa * (dot(e.mu, w) - 1));
    return e;
}

// r: lobe k's raw outputs; pi = its softmax weight; q = the mixture density
// at w (at least 1e-30).
LAM_FN void guide_lobe_grad(const float *r, V3 w, float c, const LobeEval &e, float pi, float q, float *g) {
    const float resp = pi * e.comp / q;  // posterior of lobe k
    // d(-c log q)/d logit_k = -c (resp_k - pi_k)
    g[4] = -c * (resp - pi);
    // d log vMF / d kappa = 1/kappa - 2 exp(-2k)/(1 - exp(-2k)) + (mu.w - 1)
    const float kap = e.kappa;
    const float ex = expf(-2 * kap);
    const float dlog_dk = 1 / kap - 2 * ex / (-expm1f(-2 * kap)) + (dot(e.mu, w) - 1);
    const float ek = expf(r[3]);
    const float dk_draw = (ek > kKappaMin && ek < kKappaMax) ? ek : 0.0f;
    g[3] = -c * resp * dlog_dk * dk_draw;
    // d log vMF / d mu = kappa w; mu = m / |m|: d/dm = (I - mu mu^T) / |m|
    V3 m = v3(r[0], r[1], r[2]);
    const float l = fmaxf(len(m), 1e-6f);
    V3 gm = w * kap;
    gm = (gm - e.mu * dot(e.mu, gm)) * (1 / l);
    g[0] = -c * resp * gm.x;
    g[1] = -c * resp * gm.y;
    g[2] = -c * resp * gm.z;
}

// The whole row: loss and gradient with respect to the 40 raw outputs
// (d[40]); returns the loss.
LAM_FN float guide_loss_grad(const float *o, V3 w, float c, float *d) {
    LobeEval e[kLobes];
    float mx = -INFINITY;
    for (int k = 0; k < kLobes; k++) mx = fmaxf(mx, o[5 * k + 4]);
    float pi[kLobes], sum = 0;
    for (int k = 0; k < kLobes; k++) {
        e[k] = guide_lobe_density(o + 5 * k, w);
        pi[k] = expf(o[5 * k + 4] - mx);
        sum += pi[k];
    }
    float q = 0;
    for (int k = 0; k < kLobes; k++) {
        pi[k] /= sum;
        q += pi[k] * e[k].comp;
    }
    q = fmaxf(q, 1e-30f);
    for (int k = 0; k < kLobes; k++) guide_lobe_grad(o + 5 * k, w, c, e[k], pi[k], q, d + 5 * k);
    return -c * logf(q);
}

}  // namespace lam
