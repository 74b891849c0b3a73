// Device-side arithmetic for the gfx950 integrator.
//
// Vector, PCG32, sampling, transform and noise helpers.  Geometry code keeps
// Lambda's float operation order (math/vector.h, core/utils.h) and the
// library is built with -ffp-contract=off, so ray/box and ray/primitive tests
// round exactly like the reference CPU build (which has no FMA) and hit IDs
// match bit for bit.  Transcendentals (sin, cos, pow, atan2, acos) come from
// the device math library and may differ from glibc in the last place.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Callable from host code too, so the device arithmetic can be exercised on
// the CPU by test programs.
#define LAM_FN __host__ __device__ __forceinline__
// Heavy functions that few vertices reach (thin-film transfer matrices,
// procedural textures): kept out of line, so that each kernel carries one
// copy instead of one per call site (the persistent kernel's code shrinks
// from 546 KB to 100 KB plus 56 KB for the two callees; 1080p frame
// 2.5 % faster).
#define LAM_OUTLINE inline __host__ __device__ __noinline__

namespace lam {

constexpr float kEps = float(1e-4);                       // core/constants.h:5
constexpr float kEps2 = kEps * kEps;                      // core/constants.h:6
constexpr float kPi = float(3.14159265358979323846);      // core/constants.h:7
constexpr float kMax = float(1e30);                       // core/constants.h:16
constexpr int kHero = 4;                                  // core/constants.h:12
constexpr int kRRStart = 5;                               // core/constants.h:14
constexpr int kBins = 471;
constexpr int kStack = 20;                                // core/constants.h:10

struct V3 {
    float x, y, z;
};

// Bit casts usable on host and device.
LAM_FN int as_int(float x) { return __builtin_bit_cast(int, x); }
LAM_FN float as_float(int x) { return __builtin_bit_cast(float, x); }

LAM_FN V3 v3(float a, float b, float c) { return V3{a, b, c}; }
LAM_FN V3 v3(float4 a) { return V3{a.x, a.y, a.z}; }
LAM_FN V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
LAM_FN V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
LAM_FN V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
LAM_FN V3 operator*(V3 a, float d) { return v3(a.x * d, a.y * d, a.z * d); }
LAM_FN V3 operator*(float d, V3 a) { return a * d; }
LAM_FN V3 operator/(V3 a, float d) { return v3(a.x / d, a.y / d, a.z / d); }
LAM_FN float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LAM_FN float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
LAM_FN float len(V3 a) { return sqrtf(len2(a)); }
LAM_FN V3 normalize(V3 a) {
    float l = len(a);
    return v3(a.x / l, a.y / l, a.z / l);
}
LAM_FN V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
LAM_FN float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Four hero-wavelength samples (math/spectrum.h:8-156).
struct S4 {
    float v[kHero];
};
LAM_FN S4 s4(float d) { return S4{{d, d, d, d}}; }
LAM_FN S4 s4(float4 a) { return S4{{a.x, a.y, a.z, a.w}}; }
LAM_FN float4 f4(const S4 &a) { return make_float4(a.v[0], a.v[1], a.v[2], a.v[3]); }
LAM_FN S4 operator*(const S4 &a, const S4 &b) {
    S4 r;
#pragma unroll
    for (int i = 0; i < kHero; i++) r.v[i] = a.v[i] * b.v[i];
    return r;
}
LAM_FN S4 operator*(const S4 &a, float d) {
    S4 r;
#pragma unroll
    for (int i = 0; i < kHero; i++) r.v[i] = a.v[i] * d;
    return r;
}
LAM_FN S4 operator/(const S4 &a, float d) {
    S4 r;
#pragma unroll
    for (int i = 0; i < kHero; i++) r.v[i] = a.v[i] / d;
    return r;
}
LAM_FN float s4max(const S4 &a) {
    float m = a.v[0];
#pragma unroll
    for (int i = 1; i < kHero; i++)
        if (a.v[i] > m) m = a.v[i];
    return m;
}

// PCG32 XSH-RR (O'Neill); seeding as pcg32_srandom_r (math/random.h:9-31).
struct Pcg {
    uint64_t state, inc;
    LAM_FN uint32_t step() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    // u32 / 2^32 rounded to float; may round up to 1.0 like the reference.
    LAM_FN float next() { return float(step()) / float(4294967296.0); }
};
LAM_FN Pcg pcg_seed(uint64_t seed, uint64_t stream) {
    Pcg r;
    r.state = 0;
    r.inc = (stream << 1u) | 1u;
    r.step();
    r.state += seed;
    r.step();
    return r;
}

LAM_FN float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Normalised uniform cube sample; the three draws land in z, y, x in that
// order, as the reference's compiled argument evaluation does
// (core/utils.h:20-24).
LAM_FN V3 random_unit_vector(Pcg &r) {
    float c = r.next() * 2 - 1;
    float b = r.next() * 2 - 1;
    float a = r.next() * 2 - 1;
    return normalize(v3(a, b, c));
}

LAM_FN void onb(V3 n, V3 &u, V3 &v, V3 &w) {
    w = normalize(n);
    u = fabsf(w.x) > fabsf(w.y) ? v3(0, 0, 1) : v3(1, 0, 0);
    u = normalize(cross(u, w));
    v = cross(w, u);
}

// Cosine-weighted direction about n (core/utils.h:26-44).
LAM_FN V3 random_in_hemisphere(V3 n, Pcg &r) {
    float r1 = r.next();
    float r2 = r.next();
    float phi = 2 * kPi * r1;
    float rr = sqrtf(r2);
    float x = rr * cosf(phi);
    float y = rr * sinf(phi);
    float z = sqrtf(1 - r2);
    V3 u, v, w;
    onb(n, u, v, w);
    return x * u + y * v + z * w;
}

// Uniform direction in the cone of half-angle acos(cos_max) about n
// (core/utils.h:46-65).
LAM_FN V3 random_in_cone(V3 n, float cos_max, Pcg &r) {
    float r1 = r.next();
    float r2 = r.next();
    float ct = 1 + r1 * (cos_max - 1);
    float st = sqrtf(1 - ct * ct);
    float phi = 2 * kPi * r2;
    float x = st * cosf(phi);
    float y = st * sinf(phi);
    V3 u, v, w;
    onb(n, u, v, w);
    return x * u + y * v + ct * w;
}

LAM_FN float power_heuristic(float a, float b) {
    a *= a;
    b *= b;
    return a / (a + b);
}

LAM_FN V3 reflected(V3 v, V3 n) { return v - 2 * dot(v, n) * n; }

LAM_FN V3 refracted(V3 v, V3 n, float ratio) {
    float vn = dot(v, n);
    float root = 1 - ratio * ratio * (1 - vn * vn);
    if (root < 0) return v - 2 * vn * n;
    return ratio * v - (ratio * vn + sqrtf(root)) * n;
}

// Row-major 4x4 homogeneous transform with the divide by the fourth row,
// replaced by 1 when it is below 1e-4 in magnitude (core/utils.h:89-95).
// rows[k] holds matrix row k.
LAM_FN V3 xform(const float4 *rows, V3 v, float w) {
    float4 r0 = rows[0], r1 = rows[1], r2 = rows[2], r3 = rows[3];
    float den = r3.x * v.x + r3.y * v.y + r3.z * v.z + r3.w * w;
    if (fabsf(den) < kEps) den = 1;
    V3 o = v3(r0.x * v.x + r0.y * v.y + r0.z * v.z + r0.w * w, r1.x * v.x + r1.y * v.y + r1.z * v.z + r1.w * w,
              r2.x * v.x + r2.y * v.y + r2.z * v.z + r2.w * w);
    return o / den;
}
// The same with the transposed matrix: column k plays the role of row k
// (normals use the transposed inverse, scene/instance.h:48).
LAM_FN V3 xform_t(const float4 *rows, V3 v, float w) {
    float4 r0 = rows[0], r1 = rows[1], r2 = rows[2], r3 = rows[3];
    float den = r0.w * v.x + r1.w * v.y + r2.w * v.z + r3.w * w;
    if (fabsf(den) < kEps) den = 1;
    V3 o = v3(r0.x * v.x + r1.x * v.y + r2.x * v.z + r3.x * w, r0.y * v.x + r1.y * v.y + r2.y * v.z + r3.y * w,
              r0.z * v.x + r1.z * v.y + r2.z * v.z + r3.z * w);
    return o / den;
}

// ---------------------------------------------------------------- noise
// Improved Perlin noise remapped to [0,1] (core/utils.h:97-148).
LAM_FN float fade(float t) { return powf(t, 3.0f) * (t * (6 * t - 15) + 10); }
LAM_FN float grad(int h, float x, float y, float z) {
    switch (h % 12) {
        case 0: return x + y;
        case 1: return -x + y;
        case 2: return x - y;
        case 3: return -x - y;
        case 4: return x + z;
        case 5: return -x + z;
        case 6: return x - z;
        case 7: return -x - z;
        case 8: return y + z;
        case 9: return -y + z;
        case 10: return y - z;
        default: return -y - z;
    }
}
LAM_FN float lerp(float a, float b, float x) { return a + x * (b - a); }

LAM_FN float perlin(const int *perm, V3 p) {
    float fx = floorf(p.x), fy = floorf(p.y), fz = floorf(p.z);
    int i = (int)fx & 255, j = (int)fy & 255, k = (int)fz & 255;
    float tx = p.x - fx, ty = p.y - fy, tz = p.z - fz;
    float u = fade(tx), v = fade(ty), w = fade(tz);
    auto P = [&](int q) { return perm[q & 255]; };
    int pi0 = P(i), pi1 = P(i + 1);
    int a0 = P(pi0 + j), a1 = P(pi0 + j + 1), b0 = P(pi1 + j), b1 = P(pi1 + j + 1);
    int aaa = P(a0 + k), aba = P(a1 + k), aab = P(a0 + k + 1), abb = P(a1 + k + 1);
    int baa = P(b0 + k), bba = P(b1 + k), bab = P(b0 + k + 1), bbb = P(b1 + k + 1);
    float x1 = lerp(grad(aaa, tx, ty, tz), grad(baa, tx - 1, ty, tz), u);
    float x2 = lerp(grad(aba, tx, ty - 1, tz), grad(bba, tx - 1, ty - 1, tz), u);
    float y1 = lerp(x1, x2, v);
    x1 = lerp(grad(aab, tx, ty, tz - 1), grad(bab, tx - 1, ty, tz - 1), u);
    x2 = lerp(grad(abb, tx, ty - 1, tz - 1), grad(bbb, tx - 1, ty - 1, tz - 1), u);
    float y2 = lerp(x1, x2, v);
    return (lerp(y1, y2, w) + 1) / 2;
}

LAM_FN float fract(float x) { return x - floorf(x); }

// Feature point of a unit cell: the cell corner plus a hash of its integer
// coordinates, fract(sin(dot(cell, k)) * 43758.5453) per axis
// (core/utils.h:161-166).
LAM_FN V3 worley_feature(int cx, int cy, int cz) {
    V3 c = v3(float(cx), float(cy), float(cz));
    const V3 k1 = v3(float(127.1), float(311.7), float(74.7));
    const V3 k2 = v3(float(269.5), float(183.3), float(246.1));
    const V3 k3 = v3(float(113.5), float(271.9), float(124.6));
    const float amp = float(43758.5453);
    V3 f = v3(fract(sinf(dot(c, k1)) * amp), fract(sinf(dot(c, k2)) * amp), fract(sinf(dot(c, k3)) * amp));
    return f + c;
}

// Precomputed feature points of a box of cells.  The hash multiplies the
// last bit of sin() by 43758, so the points are computed once on the host,
// with the same float operations and the host's libm, and looked up here;
// cells outside the box are hashed on the device.
struct WorleyTable {
    const float4 *points;  // [nz][ny][nx] feature points
    int x0, y0, z0, nx, ny, nz;
};

// Cellular (F1) distance to one feature point per unit cell over the 27
// cells around p (core/utils.h:150-171).
LAM_FN float worley(V3 p, const WorleyTable &tab) {
    int xi = int(floorf(p.x)), yi = int(floorf(p.y)), zi = int(floorf(p.z));
    float d = 9999;
    for (int xo = -1; xo <= 1; xo++)
        for (int yo = -1; yo <= 1; yo++)
            for (int zo = -1; zo <= 1; zo++) {
                int cx = xi + xo, cy = yi + yo, cz = zi + zo;
                unsigned lx = unsigned(cx - tab.x0), ly = unsigned(cy - tab.y0), lz = unsigned(cz - tab.z0);
                V3 f;
                if (tab.points && lx < unsigned(tab.nx) && ly < unsigned(tab.ny) && lz < unsigned(tab.nz))
                    f = v3(tab.points[(size_t(lz) * tab.ny + ly) * tab.nx + lx]);
                else
                    f = worley_feature(cx, cy, cz);
                d = fminf(d, len(p - f));
            }
    return d;
}

}  // namespace lam
