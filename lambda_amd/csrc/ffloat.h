// Float-float ("double-single") arithmetic: a value is an unevaluated sum
// hi + lo of two FP32 numbers with |lo| <= ulp(hi)/2, about 48 significant
// bits, built from FP32 adds, multiplies and fused multiply-adds
// (error-free transformations: Knuth's TwoSum, Dekker/FMA TwoProduct; see
// Joldes, Muller, Popescu, "Tight and rigorous error bounds for basic
// building blocks of double-word arithmetic", ACM TOMS 2017).
//
// Used where FP32 is too short but the hardware FP64 path is reserved for the
// framebuffer: the phase and admittance recursion of thin films near grazing
// incidence, where the reflectance is a resonance whose width scales with
// cos(theta).  Requires -ffp-contract=off so that the compiler does not fuse
// the error terms away.
#pragma once

#include "dmath.h"

namespace lam {

struct FF {
    float hi, lo;
};

LAM_FN FF ff(float a) { return FF{a, 0.0f}; }

LAM_FN FF quick_two_sum(float a, float b) {
    float s = a + b;
    return FF{s, b - (s - a)};
}
LAM_FN FF two_sum(float a, float b) {
    float s = a + b;
    float bb = s - a;
    return FF{s, (a - (s - bb)) + (b - bb)};
}
LAM_FN FF two_prod(float a, float b) {
    float p = a * b;
    return FF{p, fmaf(a, b, -p)};
}

LAM_FN FF operator+(FF a, FF b) {
    FF s = two_sum(a.hi, b.hi);
    FF t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}
LAM_FN FF operator-(FF a) { return FF{-a.hi, -a.lo}; }
LAM_FN FF operator-(FF a, FF b) { return a + (-b); }
LAM_FN FF operator*(FF a, FF b) {
    FF p = two_prod(a.hi, b.hi);
    p.lo = fmaf(a.hi, b.lo, p.lo);
    p.lo = fmaf(a.lo, b.hi, p.lo);
    return quick_two_sum(p.hi, p.lo);
}
LAM_FN FF operator*(FF a, float b) {
    FF p = two_prod(a.hi, b);
    p.lo = fmaf(a.lo, b, p.lo);
    return quick_two_sum(p.hi, p.lo);
}
// Reciprocal and square root to about 1 ulp (the hardware's approximate
// instructions on the device): enough for the leading quotient or root of
// the float-float division and square root, whose residual steps correct
// it, and about a fifth of the instructions of a correctly rounded FP32
// division or square root.
LAM_FN float approx_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
LAM_FN float approx_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}

// Division by Newton-style correction: three partial quotients, each from
// the residual of the previous ones.
LAM_FN FF operator/(FF a, FF b) {
    const float rb = approx_rcp(b.hi);
    float q1 = a.hi * rb;
    FF r = a - b * q1;
    float q2 = r.hi * rb;
    r = r - b * q2;
    float q3 = r.hi * rb;
    FF q = quick_two_sum(q1, q2);
    return q + ff(q3);
}
LAM_FN FF ffsqrt(FF a) {
    if (a.hi <= 0) return ff(0);
    float x = approx_sqrt(a.hi);
    // one Newton step: x + (a - x^2) / (2x)
    FF r = a - two_prod(x, x);
    return quick_two_sum(x, r.hi * (0.5f * approx_rcp(x)));
}
LAM_FN float ffabs_hi(FF a) { return fabsf(a.hi); }
LAM_FN FF ffabs(FF a) { return a.hi < 0 ? -a : a; }
LAM_FN float ffloat(FF a) { return a.hi + a.lo; }

// Terms of the sine and cosine series after the first: on |r| <= pi/4 the
// first term left out (r^15/15!, r^14/14!) is below 1e-11, far under what
// the reflectance needs (1e-5; the errors against the FP64 oracle did not
// change between 5 and 9 terms).
constexpr int kSinCosTerms = 6;
// sin and cos of a float-float argument.  Reduction by pi/2 with a
// three-part constant, then Taylor series in float-float on |r| <= pi/4.
LAM_FN void ffsincos(FF x, FF &s, FF &c) {
    const float kTwoOverPi = 0.636619772f;
    // pi/2 = P1 + P2 + P3 + O(1e-23); each k * Pi is formed exactly by
    // TwoProduct.
    const float P1 = 0x1.921fb6p+0f;
    const float P2 = -0x1.777a5cp-25f;
    const float P3 = -0x1.ee59dap-50f;
    float kf = rintf(x.hi * kTwoOverPi);
    int k = int(kf);
    FF r = x - two_prod(kf, P1);
    r = r - two_prod(kf, P2);
    r = r - two_prod(kf, P3);
    FF r2 = r * r;
    // sin r = r (1 - r^2/3! + r^4/5! - ...), cos r = 1 - r^2/2! + r^4/4! - ...
    FF ts = ff(1.0f), tc = ff(1.0f);
    FF sn = ff(1.0f), cs = ff(1.0f);
    for (int i = 1; i <= kSinCosTerms; i++) {
        ts = ts * r2 * (-1.0f / float((2 * i) * (2 * i + 1)));
        tc = tc * r2 * (-1.0f / float((2 * i - 1) * (2 * i)));
        sn = sn + ts;
        cs = cs + tc;
    }
    sn = sn * r;
    switch (k & 3) {
        case 0: s = sn; c = cs; break;
        case 1: s = cs; c = -sn; break;
        case 2: s = -sn; c = -cs; break;
        default: s = -cs; c = sn; break;
    }
}

// exp of a float-float argument: x = k ln2 + r with a three-part ln2, then a
// Taylor series for exp(r), |r| <= ln2 / 2.
LAM_FN FF ffexp(FF x) {
    if (x.hi > 88.0f) return ff(INFINITY);
    if (x.hi < -103.0f) return ff(0.0f);
    const float L1 = 0x1.62e430p-1f;
    const float L2 = -0x1.05c610p-29f;
    const float L3 = -0x1.950d88p-54f;
    float kf = rintf(x.hi * 1.44269504f);
    FF r = x - two_prod(kf, L1);
    r = r - two_prod(kf, L2);
    r = r - two_prod(kf, L3);
    FF term = ff(1.0f), sum = ff(1.0f);
    for (int i = 1; i <= 12; i++) {
        term = term * r * (1.0f / float(i));
        sum = sum + term;
    }
    float scale = ldexpf(1.0f, int(kf));
    return FF{sum.hi * scale, sum.lo * scale};
}

// tanh of a float-float argument from exp(-2|x|).
LAM_FN FF fftanh(FF x) {
    bool neg = x.hi < 0;
    FF ax = neg ? -x : x;
    FF e = ffexp(ax * -2.0f);
    FF t = (ff(1.0f) - e) / (ff(1.0f) + e);
    return neg ? -t : t;
}

// Complex float-float.
struct CFF {
    FF re, im;
};
LAM_FN CFF cff(FF r, FF i) { return CFF{r, i}; }
LAM_FN CFF cff(float r, float i = 0) { return CFF{ff(r), ff(i)}; }
LAM_FN CFF operator+(CFF a, CFF b) { return CFF{a.re + b.re, a.im + b.im}; }
LAM_FN CFF operator-(CFF a, CFF b) { return CFF{a.re - b.re, a.im - b.im}; }
LAM_FN CFF operator*(CFF a, CFF b) { return CFF{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
LAM_FN CFF operator*(CFF a, FF s) { return CFF{a.re * s, a.im * s}; }
LAM_FN CFF operator/(CFF a, CFF b) {
    // Smith's scaling on the leading parts.
    if (fabsf(b.re.hi) >= fabsf(b.im.hi)) {
        FF r = b.im / b.re;
        FF den = b.re + b.im * r;
        return CFF{(a.re + a.im * r) / den, (a.im - a.re * r) / den};
    }
    FF r = b.re / b.im;
    FF den = b.re * r + b.im;
    return CFF{(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}
LAM_FN FF cffabs2(CFF a) { return a.re * a.re + a.im * a.im; }
// Principal square root with the sign of a zero imaginary part selecting the
// branch on the negative real axis.
LAM_FN CFF cffsqrt(CFF z) {
    FF m = ffsqrt(cffabs2(z));
    if (m.hi == 0) return CFF{ff(0), z.im};
    if (z.re.hi >= 0) {
        FF t = ffsqrt((m + z.re) * 0.5f);
        return CFF{t, z.im / (t * 2.0f)};
    }
    FF t = ffsqrt((m - z.re) * 0.5f);
    FF re = ffabs(z.im) / (t * 2.0f);
    return CFF{re, copysignf(1.0f, z.im.hi) < 0 ? -t : t};
}

}  // namespace lam
