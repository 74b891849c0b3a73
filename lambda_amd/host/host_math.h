// Host-side vector and matrix arithmetic for the scene front end.
//
// Every operation evaluates in float with the same operand order as Lambda's
// value types (math/vector.h, math/matrix.h, core/utils.h:89-95), so that the
// bounds, centroids, transforms and light powers the front end derives are
// bit-identical to the reference's.  Build without FMA contraction
// (-ffp-contract=off, no -march): a fused multiply-add rounds once where the
// reference rounds twice.
#pragma once

#include <cmath>

namespace lambda_host {

constexpr float kEpsilon = float(1e-4);                       // core/constants.h:5
constexpr float kPi = float(3.14159265358979323846);    // core/constants.h:7
constexpr float kMax = float(1e30);                           // core/constants.h:16
constexpr int kBinCount = 16;                           // core/constants.h:9
constexpr int kMaxDepth = 20;                           // core/constants.h:10
constexpr int kBins = 471;                              // core/constants.h:22
constexpr int kLambdaMin = 360;
constexpr int kLambdaMax = 830;

struct Vec3 {
    float x = 0, y = 0, z = 0;
    Vec3() = default;
    Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

inline Vec3 operator+(const Vec3 &a, const Vec3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(const Vec3 &a, const Vec3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(const Vec3 &a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator*(float s, const Vec3 &a) { return a * s; }
inline Vec3 operator/(const Vec3 &a, float s) { return {a.x / s, a.y / s, a.z / s}; }

inline float dot(const Vec3 &a, const Vec3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length_squared(const Vec3 &a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline float length(const Vec3 &a) { return std::sqrt(length_squared(a)); }
inline Vec3 normalize(const Vec3 &a) {
    float len = length(a);
    return {a.x / len, a.y / len, a.z / len};
}
inline Vec3 cross(const Vec3 &a, const Vec3 &b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline Vec3 vmin(const Vec3 &a, const Vec3 &b) {
    return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)};
}
inline Vec3 vmax(const Vec3 &a, const Vec3 &b) {
    return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)};
}

// Row-major 4x4 float matrix.
struct Mat4 {
    float m[16] = {};
    float operator()(int r, int c) const { return m[r * 4 + c]; }
};

inline Mat4 mat4(float a0, float a1, float a2, float a3, float a4, float a5, float a6, float a7,
                 float a8, float a9, float a10, float a11, float a12, float a13, float a14,
                 float a15) {
    Mat4 r;
    const float v[16] = {a0, a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11, a12, a13, a14, a15};
    for (int i = 0; i < 16; i++) r.m[i] = v[i];
    return r;
}

// Product summed left to right over k (math/matrix.h:183-200).
inline Mat4 operator*(const Mat4 &a, const Mat4 &b) {
    Mat4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r.m[i * 4 + j] = a.m[i * 4 + 0] * b.m[0 * 4 + j] + a.m[i * 4 + 1] * b.m[1 * 4 + j] +
                             a.m[i * 4 + 2] * b.m[2 * 4 + j] + a.m[i * 4 + 3] * b.m[3 * 4 + j];
    return r;
}

// Laplace expansion along the first row (math/matrix.h:243).
inline float determinant(const Mat4 &M) {
    const float *d = M.m;
    return d[0] * (d[5] * (d[10] * d[15] - d[11] * d[14]) - d[6] * (d[9] * d[15] - d[11] * d[13]) +
                   d[7] * (d[9] * d[14] - d[10] * d[13])) -
           d[1] * (d[4] * (d[10] * d[15] - d[11] * d[14]) - d[6] * (d[8] * d[15] - d[11] * d[12]) +
                   d[7] * (d[8] * d[14] - d[10] * d[12])) +
           d[2] * (d[4] * (d[9] * d[15] - d[11] * d[13]) - d[5] * (d[8] * d[15] - d[11] * d[12]) +
                   d[7] * (d[8] * d[13] - d[9] * d[12])) -
           d[3] * (d[4] * (d[9] * d[14] - d[10] * d[13]) - d[5] * (d[8] * d[14] - d[10] * d[12]) +
                   d[6] * (d[8] * d[13] - d[9] * d[12]));
}

// Inverse by 2x2 sub-determinants of the top and bottom row pairs
// (math/matrix.h:247-269); each entry is divided by the full determinant.
inline Mat4 inverse(const Mat4 &M) {
    const float *m = M.m;
    float det = determinant(M);
    float s0 = m[0] * m[5] - m[1] * m[4];
    float s1 = m[0] * m[6] - m[2] * m[4];
    float s2 = m[0] * m[7] - m[3] * m[4];
    float s3 = m[1] * m[6] - m[2] * m[5];
    float s4 = m[1] * m[7] - m[3] * m[5];
    float s5 = m[2] * m[7] - m[3] * m[6];
    float c0 = m[8] * m[13] - m[9] * m[12];
    float c1 = m[8] * m[14] - m[10] * m[12];
    float c2 = m[8] * m[15] - m[11] * m[12];
    float c3 = m[9] * m[14] - m[10] * m[13];
    float c4 = m[9] * m[15] - m[11] * m[13];
    float c5 = m[10] * m[15] - m[11] * m[14];
    return mat4((m[5] * c5 - m[6] * c4 + m[7] * c3) / det, (-m[1] * c5 + m[2] * c4 - m[3] * c3) / det,
                (m[13] * s5 - m[14] * s4 + m[15] * s3) / det, (-m[9] * s5 + m[10] * s4 - m[11] * s3) / det,
                (-m[4] * c5 + m[6] * c2 - m[7] * c1) / det, (m[0] * c5 - m[2] * c2 + m[3] * c1) / det,
                (-m[12] * s5 + m[14] * s2 - m[15] * s1) / det, (m[8] * s5 - m[10] * s2 + m[11] * s1) / det,
                (m[4] * c4 - m[5] * c2 + m[7] * c0) / det, (-m[0] * c4 + m[1] * c2 - m[3] * c0) / det,
                (m[12] * s4 - m[13] * s2 + m[15] * s0) / det, (-m[8] * s4 + m[9] * s2 - m[11] * s0) / det,
                (-m[4] * c3 + m[5] * c1 - m[6] * c0) / det, (m[0] * c3 - m[1] * c1 + m[2] * c0) / det,
                (-m[12] * s3 + m[13] * s1 - m[14] * s0) / det, (m[8] * s3 - m[9] * s1 + m[10] * s0) / det);
}

// Homogeneous transform with the divide by the fourth row; a near-zero
// divisor is replaced by 1 (core/utils.h:89-95).
inline Vec3 transform(const Mat4 &M, const Vec3 &v, float w) {
    float den = M(3, 0) * v.x + M(3, 1) * v.y + M(3, 2) * v.z + M(3, 3) * w;
    if (std::fabs(den) < kEpsilon) den = 1;
    Vec3 r(M(0, 0) * v.x + M(0, 1) * v.y + M(0, 2) * v.z + M(0, 3) * w,
           M(1, 0) * v.x + M(1, 1) * v.y + M(1, 2) * v.z + M(1, 3) * w,
           M(2, 0) * v.x + M(2, 1) * v.y + M(2, 2) * v.z + M(2, 3) * w);
    return r / den;
}

}  // namespace lambda_host
