// Test infrastructure only: CPU oracle of the spectral path tracer.
//
// A restatement of Lambda's CPU integrator over lambda_scene_view.  Each
// function cites the reference code it follows; the float operations are
// written in the reference's evaluation order so that, built like the
// reference (g++ -O3, no -march, so no FMA), the images agree bit for bit.
// The product never links this file.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

#include "spectral_tables.inc"

using Complex = std::complex<double>;

constexpr float EPS = float(1e-4);                          // core/constants.h:5
constexpr float EPS2 = EPS * EPS;                     // core/constants.h:6
constexpr float PI = float(3.14159265358979323846);   // core/constants.h:7
constexpr float MAXF = float(1e30);                         // core/constants.h:16
constexpr int HERO = 4;                               // core/constants.h:12
constexpr int RR_START = 5;                           // core/constants.h:14
constexpr int BINS = 471;
constexpr int STACK = 20;                             // core/constants.h:10

// ------------------------------------------------------------ math/vector.h
struct V3 {
    float x, y, z;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 mk(float a, float b, float c) { return V3{a, b, c}; }
inline V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return mk(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, float d) { return mk(a.x * d, a.y * d, a.z * d); }
inline V3 operator*(float d, V3 a) { return a * d; }
inline V3 operator/(V3 a, float d) { return mk(a.x / d, a.y / d, a.z / d); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline float len(V3 a) { return sqrtf(len2(a)); }
inline V3 normalize(V3 a) {
    float l = len(a);
    return mk(a.x / l, a.y / l, a.z / l);
}
inline V3 cross(V3 a, V3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// ---------------------------------------------------------- math/spectrum.h
struct Spec {
    float v[HERO];
    float &operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
};
inline Spec spec(float d) { return Spec{{d, d, d, d}}; }
inline Spec operator*(const Spec &a, const Spec &b) {
    Spec r;
    for (int i = 0; i < HERO; i++) r.v[i] = a.v[i] * b.v[i];
    return r;
}
inline Spec operator*(const Spec &a, float d) {
    Spec r;
    for (int i = 0; i < HERO; i++) r.v[i] = a.v[i] * d;
    return r;
}
inline Spec operator/(const Spec &a, float d) {
    Spec r;
    for (int i = 0; i < HERO; i++) r.v[i] = a.v[i] / d;
    return r;
}
inline void operator+=(Spec &a, const Spec &b) {
    for (int i = 0; i < HERO; i++) a.v[i] += b.v[i];
}
inline void operator*=(Spec &a, const Spec &b) {
    for (int i = 0; i < HERO; i++) a.v[i] *= b.v[i];
}
inline void operator/=(Spec &a, float d) {
    for (int i = 0; i < HERO; i++) a.v[i] /= d;
}
inline float smax(const Spec &a) {
    float m = a.v[0];
    for (int i = 1; i < HERO; i++)
        if (a.v[i] > m) m = a.v[i];
    return m;
}

// Dense-spectrum lookup: linear interpolation with a double-precision index
// (math/spectrum.h:170-188).  The index is the float difference lambda - 360.
inline float dense(const float *s, float lambda) {
    double index = lambda - 360;
    int lo = int(index);
    if (lo == BINS - 1) return s[lo];
    int hi = lo + 1;
    return float((hi - index) * s[lo] + (index - lo) * s[hi]);
}
inline Complex dense_c(const Complex *s, float lambda) {
    double index = lambda - 360;
    int lo = int(index);
    if (lo == BINS - 1) return s[lo];
    int hi = lo + 1;
    return Complex((hi - index) * s[lo] + (index - lo) * s[hi]);
}

// ------------------------------------------------------------ math/random.h
struct Rng {
    uint64_t state, inc;
    Rng(uint64_t seed, uint64_t stream) : state(0), inc((stream << 1u) | 1u) {
        step();
        state += seed;
        step();
    }
    uint32_t step() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    float next() { return float(step()) / float(4294967296.0); }
};

// ------------------------------------------------------------ core/utils.h
inline float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// The reference builds the vector from three draws inside one constructor
// call; g++ evaluates those arguments last to first (core/utils.h:20-24).
inline V3 random_unit_vector(Rng &r) {
    float c = r.next() * 2 - 1;
    float b = r.next() * 2 - 1;
    float a = r.next() * 2 - 1;
    return normalize(mk(a, b, c));
}

inline void frame(V3 n, V3 &u, V3 &v, V3 &w) {
    w = normalize(n);
    u = fabsf(w.x) > fabsf(w.y) ? mk(0, 0, 1) : mk(1, 0, 0);
    u = normalize(cross(u, w));
    v = cross(w, u);
}

// Cosine-weighted hemisphere (core/utils.h:26-44).
inline V3 random_in_hemisphere(V3 n, Rng &r) {
    float r1 = r.next();
    float r2 = r.next();
    float phi = 2 * PI * r1;
    float rr = sqrtf(r2);
    float x = rr * cosf(phi);
    float y = rr * sinf(phi);
    float z = sqrtf(1 - r2);
    V3 u, v, w;
    frame(n, u, v, w);
    return x * u + y * v + z * w;
}

// Uniform cone (core/utils.h:46-65).
inline V3 random_in_cone(V3 n, float cos_max, Rng &r) {
    float r1 = r.next();
    float r2 = r.next();
    float ct = 1 + r1 * (cos_max - 1);
    float st = sqrtf(1 - ct * ct);
    float phi = 2 * PI * r2;
    float x = st * cosf(phi);
    float y = st * sinf(phi);
    float z = ct;
    V3 u, v, w;
    frame(n, u, v, w);
    return x * u + y * v + z * w;
}

inline float power_heuristic(float a, float b) {
    a *= a;
    b *= b;
    return a / (a + b);
}

inline V3 reflected(V3 v, V3 n) { return v - 2 * dot(v, n) * n; }

inline V3 refracted(V3 v, V3 n, float ratio) {
    float vn = dot(v, n);
    float root = 1 - ratio * ratio * (1 - vn * vn);
    if (root < 0) return v - 2 * vn * n;
    return ratio * v - (ratio * vn + sqrtf(root)) * n;
}

// Homogeneous 4x4 transform (core/utils.h:89-95); m row-major.
inline V3 xform(const float *m, V3 v, float w) {
    float den = m[12] * v.x + m[13] * v.y + m[14] * v.z + m[15] * w;
    if (fabsf(den) < EPS) den = 1;
    V3 r = mk(m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3] * w, m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7] * w,
              m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11] * w);
    return r / den;
}
// Same with the transpose of m (instance.h:48 transforms normals by the
// transposed inverse).
inline V3 xform_t(const float *m, V3 v, float w) {
    float den = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * w;
    if (fabsf(den) < EPS) den = 1;
    V3 r = mk(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * w, m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * w,
              m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * w);
    return r / den;
}

// Improved noise (core/utils.h:97-148).
inline float fade(float t) { return powf(t, 3) * (t * (6 * t - 15) + 10); }
inline int perm(int i) { return kPerlinPermutation[i & 255]; }
inline float grad(int h, float x, float y, float z) {
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
        case 11: return -y - z;
        default: return 0;
    }
}
inline float lerp(float a, float b, float x) { return a + x * (b - a); }
float perlin(V3 p) {
    int i = (int)floorf(p.x) & 255;
    int j = (int)floorf(p.y) & 255;
    int k = (int)floorf(p.z) & 255;
    float tx = p.x - floorf(p.x);
    float ty = p.y - floorf(p.y);
    float tz = p.z - floorf(p.z);
    float u = fade(tx), v = fade(ty), w = fade(tz);
    int aaa = perm(perm(perm(i) + j) + k);
    int aba = perm(perm(perm(i) + j + 1) + k);
    int aab = perm(perm(perm(i) + j) + k + 1);
    int abb = perm(perm(perm(i) + j + 1) + k + 1);
    int baa = perm(perm(perm(i + 1) + j) + k);
    int bba = perm(perm(perm(i + 1) + j + 1) + k);
    int bab = perm(perm(perm(i + 1) + j) + k + 1);
    int bbb = perm(perm(perm(i + 1) + j + 1) + k + 1);
    float x1 = lerp(grad(aaa, tx, ty, tz), grad(baa, tx - 1, ty, tz), u);
    float x2 = lerp(grad(aba, tx, ty - 1, tz), grad(bba, tx - 1, ty - 1, tz), u);
    float y1 = lerp(x1, x2, v);
    x1 = lerp(grad(aab, tx, ty, tz - 1), grad(bab, tx - 1, ty, tz - 1), u);
    x2 = lerp(grad(abb, tx, ty - 1, tz - 1), grad(bbb, tx - 1, ty - 1, tz - 1), u);
    float y2 = lerp(x1, x2, v);
    return (lerp(y1, y2, w) + 1) / 2;
}

inline float fract(float x) { return x - floorf(x); }
// Cellular F1 distance over the 27 neighbouring cells (core/utils.h:150-171).
float worley(V3 p) {
    int xi = int(floorf(p.x)), yi = int(floorf(p.y)), zi = int(floorf(p.z));
    float d = 9999;
    for (int xo = -1; xo <= 1; xo++)
        for (int yo = -1; yo <= 1; yo++)
            for (int zo = -1; zo <= 1; zo++) {
                V3 c = mk(float(xi + xo), float(yi + yo), float(zi + zo));
                V3 f = mk(fract(sinf(dot(c, mk(float(127.1), float(311.7), float(74.7)))) * float(43758.5453)),
                          fract(sinf(dot(c, mk(float(269.5), float(183.3), float(246.1)))) * float(43758.5453)),
                          fract(sinf(dot(c, mk(float(113.5), float(271.9), float(124.6)))) * float(43758.5453)));
                f = f + c;
                d = fminf(d, len(p - f));
            }
    return d;
}

// CIE 1931 lookup (core/utils.h:195-203): table entries rounded to float.
inline float xyz31(float lambda, int i) {
    int lo = int(lambda);
    if (lo == 830) return float(kCie1931Xyz[lo - 360][i]);
    int hi = lo + 1;
    return (float(hi) - lambda) * float(kCie1931Xyz[lo - 360][i]) + (lambda - float(lo)) * float(kCie1931Xyz[hi - 360][i]);
}

inline V3 spectrum_to_rgb(const Spec &s, const Spec &lambdas, float range) {
    float x = 0, y = 0, z = 0;
    float weight = range / HERO;
    for (int i = 0; i < HERO; i++) {
        float lambda = lambdas[i];
        float radiance = s[i];
        x += xyz31(lambda, 0) * radiance * weight;
        y += xyz31(lambda, 1) * radiance * weight;
        z += xyz31(lambda, 2) * radiance * weight;
    }
    x /= float(106.856895);
    y /= float(106.856895);
    z /= float(106.856895);
    return mk(x * float(3.2406255) - y * float(1.5372080) - z * float(0.4986286), -x * float(0.9689307) + y * float(1.8757561) + z * float(0.0415175),
              x * float(0.0557101) - y * float(0.2040211) + z * float(1.0569959));
}

// 2x2 complex transfer matrices (math/matrix.h:8-120, core/utils.h:173-193).
struct M2 {
    Complex d[4];
};
inline M2 mul(const M2 &a, const M2 &b) {
    M2 r;
    r.d[0] = a.d[0] * b.d[0] + a.d[1] * b.d[2];
    r.d[1] = a.d[0] * b.d[1] + a.d[1] * b.d[3];
    r.d[2] = a.d[2] * b.d[0] + a.d[3] * b.d[2];
    r.d[3] = a.d[2] * b.d[1] + a.d[3] * b.d[3];
    return r;
}
inline M2 interface_s(Complex n1, Complex n2, Complex c1, Complex c2) {
    Complex r = (n1 * c1 - n2 * c2) / (n1 * c1 + n2 * c2);
    M2 m{{Complex(1), r, r, Complex(1)}};
    Complex t = Complex(2) * n1 * c1 / (n1 * c1 + n2 * c2);
    for (auto &x : m.d) x /= t;
    return m;
}
inline M2 interface_p(Complex n1, Complex n2, Complex c1, Complex c2) {
    Complex r = (n2 * c1 - n1 * c2) / (n2 * c1 + n1 * c2);
    M2 m{{Complex(1), r, r, Complex(1)}};
    Complex t = Complex(2) * n1 * c1 / (n2 * c1 + n1 * c2);
    for (auto &x : m.d) x /= t;
    return m;
}
inline M2 propagation(Complex phi) {
    return M2{{std::exp(Complex(0, -1) * phi), Complex(0), Complex(0), std::exp(Complex(0, 1) * phi)}};
}

// ------------------------------------------------------------- scene model
struct Object {
    int type, material;
    V3 a, h, v, n;  // sphere: a = centre, radius in r
    float r;
    V3 n0, n1, n2;
    float u0, u1, u2, v0, v1, v2;
};
struct Instance {
    int object, count, node, material;
    float M[16], Minv[16];
};
struct Node {
    V3 lo, hi;
    int index, count;
};
struct Hit {
    int instance = -1, object = -1;
    float t = MAXF;
    V3 point{0, 0, 0}, normal{0, 0, 0};
    float u = 0, v = 0;
    bool front = false;
};
struct Ray {
    V3 o, d;
    Spec lambdas;
};
// 0 paths, 1 closest rays, 2 shadow rays; 3-7 closest-hit work and 8-12
// shadow work: node tests, instance entries, sphere, quad, triangle tests.
struct Counters {
    int64_t c[13] = {0};
    int base = 3;  // 3 while tracing a closest-hit ray, 8 for a shadow ray
};

struct SceneData {
    const lambda_scene_view *sv;
    std::vector<Object> objects;
    std::vector<Instance> instances;
    std::vector<Node> nodes;
    std::vector<Complex> cspec;
    V3 cam_pos, cam_corner, cam_h, cam_v;
    const float *spectra;

    explicit SceneData(const lambda_scene_view *s) : sv(s) {
        spectra = s->spectra;
        objects.resize(size_t(s->num_objects));
        for (int i = 0; i < s->num_objects; i++) {
            const float *g = s->object_geometry + 12 * size_t(i);
            const float *sh = s->object_shading + 15 * size_t(i);
            Object &o = objects[i];
            o.type = s->object_type[i];
            o.material = s->object_material[i];
            o.a = mk(g[0], g[1], g[2]);
            o.r = g[3];
            o.h = mk(g[3], g[4], g[5]);
            o.v = mk(g[6], g[7], g[8]);
            o.n = mk(g[9], g[10], g[11]);
            o.n0 = mk(sh[0], sh[1], sh[2]);
            o.n1 = mk(sh[3], sh[4], sh[5]);
            o.n2 = mk(sh[6], sh[7], sh[8]);
            o.u0 = sh[9]; o.u1 = sh[10]; o.u2 = sh[11];
            o.v0 = sh[12]; o.v1 = sh[13]; o.v2 = sh[14];
        }
        instances.resize(size_t(s->num_instances));
        for (int i = 0; i < s->num_instances; i++) {
            const int32_t *q = s->instance_ints + 4 * size_t(i);
            Instance &in = instances[i];
            in.object = q[0];
            in.count = q[1];
            in.node = q[2];
            in.material = q[3];
            std::memcpy(in.M, s->instance_matrices + 32 * size_t(i), 64);
            std::memcpy(in.Minv, s->instance_matrices + 32 * size_t(i) + 16, 64);
        }
        nodes.resize(size_t(s->num_nodes));
        for (int i = 0; i < s->num_nodes; i++) {
            const float *b = s->node_bounds + 6 * size_t(i);
            nodes[i].lo = mk(b[0], b[1], b[2]);
            nodes[i].hi = mk(b[3], b[4], b[5]);
            nodes[i].index = s->node_ints[2 * i];
            nodes[i].count = s->node_ints[2 * i + 1];
        }
        cspec.resize(size_t(s->num_complex_spectra) * BINS);
        for (size_t i = 0; i < cspec.size(); i++) cspec[i] = Complex(s->complex_spectra[2 * i], s->complex_spectra[2 * i + 1]);
        const float *c = s->camera;
        cam_pos = mk(c[0], c[1], c[2]);
        cam_corner = mk(c[3], c[4], c[5]);
        cam_h = mk(c[6], c[7], c[8]);
        cam_v = mk(c[9], c[10], c[11]);
    }

    const float *spectrum(int i) const { return spectra + size_t(i) * BINS; }
    const Complex *cspectrum(int i) const { return cspec.data() + size_t(i) * BINS; }

    // ------------------------------------------------ scene/texture.h
    float scalar_texture(int t, const Hit &h) const {
        const float *p = sv->scalar_texture_params + 3 * t;
        switch (sv->scalar_texture_type[t]) {
            case LAMBDA_SCALAR_CONSTANT: return p[0];
            case LAMBDA_SCALAR_PERLIN: return p[0] + (p[1] - p[0]) * perlin(h.point * p[2]);
            case LAMBDA_SCALAR_WORLEY: return p[0] + (p[1] - p[0]) * worley(h.point * p[2]);
        }
        return 0;
    }
    float spectrum_texture(int t, const Hit &h, float lambda) const {
        const int32_t *ix = sv->spectrum_texture_index + 2 * t;
        switch (sv->spectrum_texture_type[t]) {
            case LAMBDA_SPECTRUM_CONSTANT: return dense(spectrum(ix[0]), lambda);
            case LAMBDA_SPECTRUM_CHECKER: {
                float scale = sv->spectrum_texture_scale[t];
                int x = int(floorf(h.u / scale));
                int y = int(floorf(h.v / scale));
                return (x + y) % 2 == 0 ? dense(spectrum(ix[0]), lambda) : dense(spectrum(ix[1]), lambda);
            }
            case LAMBDA_SPECTRUM_SCALAR: return scalar_texture(ix[0], h);
        }
        return 0;
    }
    float spectrum_texture_average(int t) const {
        const int32_t *ix = sv->spectrum_texture_index + 2 * t;
        auto avg = [&](int s) {
            float sum = 0;
            for (int i = 0; i < BINS; i++) sum += spectrum(s)[i];
            return sum / BINS;
        };
        switch (sv->spectrum_texture_type[t]) {
            case LAMBDA_SPECTRUM_CONSTANT: return avg(ix[0]);
            case LAMBDA_SPECTRUM_CHECKER: return float(0.5) * (avg(ix[0]) + avg(ix[1]));
            case LAMBDA_SPECTRUM_SCALAR: {
                int s = ix[0];
                const float *p = sv->scalar_texture_params + 3 * s;
                return sv->scalar_texture_type[s] == LAMBDA_SCALAR_CONSTANT ? p[0] : (p[0] + p[1]) / 2;
            }
        }
        return 0;
    }

    // --------------------------------------------- scene/background.h
    Spec background(const Ray &r) const {
        Hit h;
        h.point = r.d;
        h.u = (atan2f(r.d.z, r.d.x) + PI) / (2 * PI);
        h.v = acosf(r.d.y) / PI;
        Spec s;
        for (int i = 0; i < HERO; i++) s[i] = spectrum_texture(sv->background_texture, h, r.lambdas[i]);
        return s;
    }

    // ------------------------------------------------ scene/object.h
    bool intersect(int i, const Ray &r, Hit &h, Counters *cnt) const {
        const Object &o = objects[i];
        if (o.type == LAMBDA_OBJECT_SPHERE) {
            if (cnt) cnt->c[cnt->base + 2]++;
            V3 oc = r.o - o.a;
            float a = len2(r.d);
            float b = 2 * dot(r.d, oc);
            float c = dot(oc, oc) - o.r * o.r;
            float d = b * b - 4 * a * c;
            if (d < 0) return false;
            d = sqrtf(d);
            float t1 = (-b - d) / (2 * a);
            float t2 = (-b + d) / (2 * a);
            if (t1 < EPS && t2 < EPS) return false;
            float t = (t1 < EPS) ? t2 : t1;
            if (t >= h.t) return false;
            h.object = i;
            h.t = t;
            h.point = r.o + t * r.d;
            h.normal = normalize(h.point - o.a);
            h.u = (atan2f(-h.normal.z, h.normal.x) + PI) / (2 * PI);
            h.v = acosf(-h.normal.y) / PI;
            h.front = dot(r.d, h.normal) < 0;
            if (!h.front) h.normal = h.normal * -1;
            return true;
        }
        if (o.type == LAMBDA_OBJECT_QUAD) {
            if (cnt) cnt->c[cnt->base + 3]++;
            V3 n = normalize(o.n);
            float den = dot(n, r.d);
            if (fabsf(den) < EPS) return false;
            float t = dot(n, o.a - r.o) / den;
            if (t < EPS || t >= h.t) return false;
            V3 p = r.o + t * r.d;
            V3 proj = p - o.a;
            float alpha = dot(o.n, cross(proj, o.v)) / len2(o.n);
            float beta = dot(o.n, cross(o.h, proj)) / len2(o.n);
            if (alpha < 0 || alpha > 1 || beta < 0 || beta > 1) return false;
            h.object = i;
            h.t = t;
            h.point = p;
            h.normal = normalize(o.n);
            h.u = alpha;
            h.v = beta;
            h.front = dot(r.d, h.normal) < 0;
            if (!h.front) h.normal = h.normal * -1;
            return true;
        }
        if (cnt) cnt->c[cnt->base + 4]++;
        V3 rcv = cross(r.d, o.v);
        float det = dot(o.h, rcv);
        if (fabsf(det) < EPS2) return false;
        float inv = 1 / det;
        V3 oc = r.o - o.a;
        float alpha = inv * dot(oc, rcv);
        if (alpha < 0 || alpha > 1) return false;
        V3 och = cross(oc, o.h);
        float beta = inv * dot(r.d, och);
        if (beta < 0 || alpha + beta > 1) return false;
        float t = inv * dot(o.v, och);
        if (t < EPS || t >= h.t) return false;
        h.object = i;
        h.t = t;
        h.point = r.o + t * r.d;
        h.normal = normalize((1 - alpha - beta) * o.n0 + alpha * o.n1 + beta * o.n2);
        h.u = (1 - alpha - beta) * o.u0 + alpha * o.u1 + beta * o.u2;
        h.v = (1 - alpha - beta) * o.v0 + alpha * o.v1 + beta * o.v2;
        h.front = dot(r.d, o.n) < 0;
        if (!h.front) h.normal = h.normal * -1;
        return true;
    }

    float object_pdf(int i, V3 p, V3 d) const {
        const Object &o = objects[i];
        if (o.type == LAMBDA_OBJECT_SPHERE) {
            V3 oc = o.a - p;
            float dist2 = len2(oc);
            float r2 = o.r * o.r;
            if (dist2 <= r2) return 1 / (4 * PI);
            float cmax = sqrtf(1 - r2 / dist2);
            float ct = dot(normalize(oc), d);
            if (ct < cmax) return 0;
            return 1 / (2 * PI * (1 - cmax));
        }
        if (o.type == LAMBDA_OBJECT_QUAD) {
            float ct = dot(normalize(o.n), d);
            if (fabsf(ct) < EPS) return 0;
            float t = dot(o.a - p, normalize(o.n)) / ct;
            if (t < EPS) return 0;
            V3 local = (p + d * t) - o.a;
            float u = dot(o.n, cross(local, o.v)) / len2(o.n);
            float v = dot(o.n, cross(o.h, local)) / len2(o.n);
            if (u < 0 || u > 1 || v < 0 || v > 1) return 0;
            return t * t / (fabsf(ct) * len(o.n));
        }
        V3 rcv = cross(d, o.v);
        float det = dot(o.h, rcv);
        if (fabsf(det) < EPS2) return 0;
        float inv = 1 / det;
        V3 oc = p - o.a;
        float alpha = inv * dot(oc, rcv);
        if (alpha < 0 || alpha > 1) return 0;
        V3 och = cross(oc, o.h);
        float beta = inv * dot(d, och);
        if (beta < 0 || alpha + beta > 1) return 0;
        float t = inv * dot(o.v, och);
        if (t < EPS) return 0;
        float area = len(o.n) * 0.5f;
        float ct = dot(normalize(o.n), d);
        return t * t / (fabsf(ct) * area);
    }

    V3 object_sample(int i, V3 p, Rng &rng) const {
        const Object &o = objects[i];
        if (o.type == LAMBDA_OBJECT_SPHERE) {
            V3 oc = o.a - p;
            float dist2 = len2(oc);
            float r2 = o.r * o.r;
            if (dist2 <= r2) return random_unit_vector(rng);
            float cmax = sqrtf(1 - r2 / dist2);
            return random_in_cone(oc, cmax, rng);
        }
        float u = rng.next();
        float v = rng.next();
        if (o.type == LAMBDA_OBJECT_TRI && u + v > 1) {
            u = 1 - u;
            v = 1 - v;
        }
        V3 s = o.a + o.h * u + o.v * v;
        return normalize(s - p);
    }

    float object_area(int i) const {
        const Object &o = objects[i];
        if (o.type == LAMBDA_OBJECT_SPHERE) return 4 * PI * o.r * o.r;
        if (o.type == LAMBDA_OBJECT_QUAD) return len(o.n);
        return len(o.n) / 2;
    }

    // ---------------------------------------------- scene/instance.h
    int material_of(const Instance &in, int rel) const {
        return in.material >= 0 ? in.material : objects[in.object + rel].material;
    }
    Ray to_local(const Instance &in, const Ray &r) const {
        return Ray{xform(in.Minv, r.o, 1), xform(in.Minv, r.d, 0), r.lambdas};
    }
    float instance_area(const Instance &in, int rel) const {
        return object_area(in.object + rel) * len2(mk(in.M[0], in.M[4], in.M[8]));
    }
    float instance_pdf(const Instance &in, int rel, V3 p, V3 d) const {
        return object_pdf(in.object + rel, xform(in.Minv, p, 1), normalize(xform(in.Minv, d, 0)));
    }
    V3 instance_sample(const Instance &in, int rel, V3 p, Rng &rng) const {
        return normalize(xform(in.M, object_sample(in.object + rel, xform(in.Minv, p, 1), rng), 0));
    }

    // --------------------------------------------------- scene/bvh.h:30-52
    static bool slab(const Node &n, const Ray &r, float &tnear) {
        float t1 = 0, t2 = MAXF;
        for (int i = 0; i < 3; i++) {
            float inv = 1 / r.d[i];
            float tmin = (n.lo[i] - r.o[i]) * inv;
            float tmax = (n.hi[i] - r.o[i]) * inv;
            if (inv < 0) std::swap(tmin, tmax);
            t1 = fmaxf(tmin, t1);
            t2 = fminf(tmax, t2);
        }
        tnear = t1;
        return t1 <= t2 && t2 >= 0;
    }

    // ------------------------------------------------ scene/scene.h:51-124
    bool hit(const Ray &r, Hit &h, int occ_inst, int occ_obj, Counters *cnt) const {
        int stack[STACK];
        int sp = 0;
        stack[sp++] = 0;
        int sub[STACK];
        int ssp = 0;
        bool found = false;
        if (occ_inst != -1) {
            const Instance &in = instances[occ_inst];
            if (cnt) cnt->c[cnt->base + 1]++;
            Ray lr = to_local(in, r);
            if (!intersect(occ_obj, lr, h, cnt)) return true;
            h.instance = occ_inst;
            h.point = xform(in.M, h.point, 1);
            h.normal = normalize(xform_t(in.Minv, h.normal, 0));
        }
        while (sp > 0) {
            int cur = stack[--sp];
            float tn;
            if (cnt) cnt->c[cnt->base + 0]++;
            if (!slab(nodes[cur], r, tn) || tn > h.t) continue;
            if (nodes[cur].count > 0) {
                int ii = nodes[cur].index;
                const Instance &in = instances[ii];
                if (cnt) cnt->c[cnt->base + 1]++;
                sub[ssp++] = in.node;
                Ray lr = to_local(in, r);
                while (ssp > 0) {
                    int sc = sub[--ssp];
                    float stn;
                    if (cnt) cnt->c[cnt->base + 0]++;
                    if (!slab(nodes[sc], lr, stn) || stn > h.t) continue;
                    if (nodes[sc].count > 0) {
                        int first = nodes[sc].index;
                        for (int k = 0; k < nodes[sc].count; k++) {
                            if (occ_inst != -1 && ii == occ_inst && first + k == occ_obj) continue;
                            if (intersect(first + k, lr, h, cnt)) {
                                h.instance = ii;
                                h.point = xform(in.M, h.point, 1);
                                h.normal = normalize(xform_t(in.Minv, h.normal, 0));
                                if (occ_obj != -2) return true;
                                found = true;
                            }
                        }
                    } else {
                        sub[ssp++] = sc + 1;
                        sub[ssp++] = nodes[sc].index;
                    }
                }
            } else {
                stack[sp++] = cur + 1;
                stack[sp++] = nodes[cur].index;
            }
        }
        return found;
    }

    // ------------------------------------------------ scene/material.h
    // A square root whose argument lies on the negative real axis up to a
    // rounding residue in its imaginary part: the branch (sign of the
    // result's imaginary part) is then decided by that residue.
    // With complex indices the imaginary part is a rounding residue of
    // either sign (or an exact zero of either sign); with real indices it is
    // an exact +0 in every implementation and the branch is determined.
    static bool noise_branch(Complex z, bool complex_indices) {
        return z.real() < 0 && std::fabs(z.imag()) <= 1e-12 * std::fabs(z.real()) &&
               (z.imag() != 0 || complex_indices);
    }

    float thinfilm_reflectance(int mat, const Hit &h, const Ray &r, float lambda, double *fp64,
                               bool *ambiguous = nullptr) const {
        const int32_t *mp = sv->material_params + 3 * mat;
        const int32_t *props = sv->material_properties;
        int layers = mp[0], n_off = mp[1], d_off = mp[2];
        int j_off = h.front ? 0 : layers + 1;
        int j_sign = h.front ? 1 : -1;
        Complex n0 = dense_c(cspectrum(props[n_off + j_off]), lambda);
        Complex n1 = dense_c(cspectrum(props[n_off + j_off + j_sign]), lambda);
        Complex cos_cur = dot(h.normal, -r.d);
        Complex sin_cur;
        Complex sin_next = n0 / n1;
        sin_next *= sin_next * (Complex(1) - cos_cur * cos_cur);
        Complex cos_next = std::sqrt(Complex(1) - sin_next);
        bool any_complex = false;
        for (int j = 0; j < layers + 2; j++)
            if (dense_c(cspectrum(props[n_off + j]), lambda).imag() != 0) any_complex = true;
        bool amb = noise_branch(Complex(1) - sin_next, any_complex);
        M2 ms = interface_s(n0, n1, cos_cur, cos_next);
        M2 mpm = interface_p(n0, n1, cos_cur, cos_next);
        for (int j = 0; j < layers; j++) {
            int d_index = h.front ? j : layers - 1 - j;
            n0 = dense_c(cspectrum(props[n_off + j_off + j_sign * (j + 1)]), lambda);
            n1 = dense_c(cspectrum(props[n_off + j_off + j_sign * (j + 2)]), lambda);
            cos_cur = cos_next;
            sin_cur = sin_next;
            sin_next = n0 / n1;
            sin_next *= sin_next * sin_cur;
            cos_next = std::sqrt(Complex(1) - sin_next);
            amb = amb || noise_branch(Complex(1) - sin_next, any_complex);
            M2 is = interface_s(n0, n1, cos_cur, cos_next);
            M2 ip = interface_p(n0, n1, cos_cur, cos_next);
            Complex phi = n0 * Complex(2 * PI / lambda) * double(scalar_texture(props[d_off + d_index], h)) * cos_cur;
            M2 pr = propagation(phi);
            ms = mul(ms, mul(pr, is));
            mpm = mul(mpm, mul(pr, ip));
        }
        if (ambiguous) *ambiguous = amb;
        if (fp64) {
            double rs = std::abs(ms.d[1] / ms.d[0]);
            double rp = std::abs(mpm.d[1] / mpm.d[0]);
            *fp64 = (rs * rs + rp * rp) * 0.5;
        }
        float rs = float(std::abs(ms.d[1] / ms.d[0]));
        float rp = float(std::abs(mpm.d[1] / mpm.d[0]));
        rs *= rs;
        rp *= rp;
        return (rs + rp) * float(0.5);
    }

    Spec evaluate_lambertian(int mat, const Hit &h, const Ray &r) const {
        if (dot(h.normal, r.d) <= 0) return spec(0);
        int tex = sv->material_params[3 * mat];
        Spec s;
        for (int j = 0; j < HERO; j++) s[j] = spectrum_texture(tex, h, r.lambdas[j]) / PI;
        return s;
    }

    Spec emission(int mat, const Hit &h, const Spec &lambdas) const {
        if (sv->material_type[mat] != LAMBDA_MATERIAL_EMISSIVE) return spec(0);
        int tex = sv->material_params[3 * mat];
        Spec s;
        for (int j = 0; j < HERO; j++) s[j] = spectrum_texture(tex, h, lambdas[j]);
        return s;
    }

    float pdf(int mat, const Hit &h, const Ray &r) const {
        if (sv->material_type[mat] != LAMBDA_MATERIAL_LAMBERTIAN) return 0;
        float ct = dot(h.normal, r.d);
        return ct > 0 ? ct / PI : 0;
    }

    // Material::scatter (scene/material.h:117-261); updates r in place.
    Spec scatter(int mat, const Hit &h, Ray &r, Rng &rng) const {
        const int32_t *mp = sv->material_params + 3 * mat;
        switch (sv->material_type[mat]) {
            case LAMBDA_MATERIAL_LAMBERTIAN: {
                r = Ray{h.point, random_in_hemisphere(h.normal, rng), r.lambdas};
                return evaluate_lambertian(mat, h, r);
            }
            case LAMBDA_MATERIAL_MIRROR: {
                r = Ray{h.point, reflected(r.d, h.normal), r.lambdas};
                Spec s;
                for (int j = 0; j < HERO; j++) s[j] = spectrum_texture(mp[0], h, r.lambdas[j]);
                return s;
            }
            case LAMBDA_MATERIAL_DIELECTRIC: {
                float factor = powf(1 + dot(r.d, h.normal), 5);
                float lambda = r.lambdas[0];
                float n0h = dense(spectrum(mp[0]), lambda);
                float n1h = dense(spectrum(mp[1]), lambda);
                float ratio = h.front ? n0h / n1h : n1h / n0h;
                float r0 = (n0h - n1h) / (n0h + n1h);
                r0 *= r0;
                float rh = r0 + (1 - r0) * factor;
                bool refl = rng.next() < rh;
                V3 dir = refl ? reflected(r.d, h.normal) : refracted(r.d, h.normal, ratio);
                Spec att;
                att[0] = 1;
                for (int j = 1; j < HERO; j++) {
                    lambda = r.lambdas[j];
                    float n0 = dense(spectrum(mp[0]), lambda);
                    float n1 = dense(spectrum(mp[1]), lambda);
                    r0 = (n0 - n1) / (n0 + n1);
                    r0 *= r0;
                    float rj = r0 + (1 - r0) * factor;
                    att[j] = refl ? rj / rh
                                  : (fabsf(n0 - n0h) < EPS && fabsf(n1 - n1h) < EPS) ? (1 - rj) / (1 - rh) : 0;
                }
                r = Ray{h.point, dir, r.lambdas};
                return att;
            }
            case LAMBDA_MATERIAL_THINFILM: {
                const int32_t *props = sv->material_properties;
                float lambda = r.lambdas[0];
                float n0h = float(dense_c(cspectrum(props[mp[1]]), lambda).real());
                float n1h = float(dense_c(cspectrum(props[mp[1] + mp[0] + 1]), lambda).real());
                float ratio = h.front ? n0h / n1h : n1h / n0h;
                float rh = thinfilm_reflectance(mat, h, r, lambda, nullptr);
                bool refl = rng.next() < rh;
                V3 dir = refl ? reflected(r.d, h.normal) : refracted(r.d, h.normal, ratio);
                Spec att;
                att[0] = 1;
                for (int j = 1; j < HERO; j++) {
                    lambda = r.lambdas[j];
                    float n0 = float(dense_c(cspectrum(props[mp[1]]), lambda).real());
                    float n1 = float(dense_c(cspectrum(props[mp[1] + mp[0] + 1]), lambda).real());
                    float rj = thinfilm_reflectance(mat, h, r, lambda, nullptr);
                    att[j] = refl ? rj / rh
                                  : (fabsf(n0 - n0h) < EPS && fabsf(n1 - n1h) < EPS) ? (1 - rj) / (1 - rh) : 0;
                }
                r = Ray{h.point, dir, r.lambdas};
                return att;
            }
        }
        return spec(0);
    }

    // --------------------------------------------- render/renderer.h:94-195
    Spec trace(Ray r, Rng &rng, int depth, Counters *cnt) const {
        Spec radiance = spec(0);
        Spec beta = spec(1);
        float prev_pdf = 1;
        bool specular = true;
        const float total = sv->total_light_power;
        const int nl = sv->num_lights;
        for (int i = 0; i <= depth; i++) {
            Hit h;
            if (cnt) {
                cnt->c[1]++;
                cnt->base = 3;
            }
            if (!hit(r, h, -1, -2, cnt)) {
                float w = 1;
                if (!specular && total > 0)
                    w = power_heuristic(prev_pdf, 1 / (4 * PI) * sv->light_powers[nl - 1] / total);
                radiance += beta * background(r) * w;
                break;
            }
            const Instance &in = instances[h.instance];
            int rel = h.object - in.object;
            int mat = material_of(in, rel);
            int mtype = sv->material_type[mat];
            if (mtype == LAMBDA_MATERIAL_EMISSIVE) {
                float w = 1;
                if (!specular && total > 0) {
                    float lp = instance_area(in, rel) * spectrum_texture_average(sv->material_params[3 * mat]);
                    w = power_heuristic(prev_pdf, instance_pdf(in, rel, r.o, r.d) * lp / total);
                }
                radiance += beta * emission(mat, h, r.lambdas) * w;
                break;
            }
            if (mtype == LAMBDA_MATERIAL_LAMBERTIAN && total > 0) {
                float target = total * rng.next();
                float cum = 0;
                int li = sv->light_instances[0], lo = sv->light_objects[0];
                float lp = sv->light_powers[0];
                for (int j = 0; j < nl; j++) {
                    cum += sv->light_powers[j];
                    if (cum >= target) {
                        li = sv->light_instances[j];
                        lo = sv->light_objects[j];
                        lp = sv->light_powers[j];
                        break;
                    }
                }
                int lrel = li != -1 ? lo - instances[li].object : -1;
                V3 ldir = li != -1 ? instance_sample(instances[li], lrel, h.point, rng) : random_unit_vector(rng);
                float lprob = li != -1 ? instance_pdf(instances[li], lrel, h.point, ldir) : 1 / (4 * PI);
                float ct = dot(h.normal, ldir);
                if (lprob > 0 && ct > 0) {
                    lprob *= lp / total;
                    Ray sr{h.point, ldir, r.lambdas};
                    Hit sh;
                    if (cnt) {
                        cnt->c[2]++;
                        cnt->base = 8;
                    }
                    if (!hit(sr, sh, li, lo, cnt)) {
                        Spec em = li != -1 ? emission(material_of(instances[li], lrel), sh, sr.lambdas) : background(sr);
                        Spec att = evaluate_lambertian(mat, h, sr);
                        float sp = pdf(mat, h, sr);
                        float w = power_heuristic(lprob, sp);
                        radiance += beta * att * ct * em * w / lprob;
                    }
                }
            }
            Spec att = scatter(mat, h, r, rng);
            prev_pdf = pdf(mat, h, r);
            specular = mtype != LAMBDA_MATERIAL_LAMBERTIAN;
            if (specular) beta *= att;
            else beta *= att * dot(h.normal, r.d) / prev_pdf;
            if (i >= RR_START) {
                float q = 1 - clampf(smax(beta), float(0.05), float(0.95));
                if (rng.next() < q) break;
                beta /= 1 - q;
            }
        }
        return radiance;
    }

    // --------------------------------------------- render/renderer.h:64-92
    V3 render_pixel(const lambda_render_params &p, int px, int py, Rng &rng, Counters *cnt) const {
        V3 color = mk(0, 0, 0);
        const float lrange = p.lambda_max - p.lambda_min;
        for (int i = 0; i < p.sqrt_samples; i++)
            for (int j = 0; j < p.sqrt_samples; j++) {
                float u = (float(px) + (float(i) + rng.next()) / float(p.sqrt_samples)) / float(p.width);
                float v = (float(py) + (float(j) + rng.next()) / float(p.sqrt_samples)) / float(p.height);
                float lambda = p.lambda_min + rng.next() * (p.lambda_max - p.lambda_min);
                Spec lambdas;
                for (int k = 0; k < HERO; k++)
                    lambdas[k] = p.lambda_min + fmodf(lambda - p.lambda_min + float(k) * lrange / HERO, lrange);
                V3 dir = normalize(cam_corner + cam_h * u + cam_v * v - cam_pos);
                Ray ray{cam_pos, dir, lambdas};
                if (cnt) cnt->c[0]++;
                Spec s = trace(ray, rng, p.depth, cnt);
                color = color + spectrum_to_rgb(s, lambdas, lrange);
            }
        return color / float(p.samples);
    }
};

bool scene_ok(const lambda_scene_view *s) {
    return s && s->num_nodes > 0 && s->num_instances > 0 && s->num_objects > 0 && s->num_lights > 0;
}

}  // namespace

extern "C" {

int oracle_render(const lambda_scene_view *scene, const lambda_render_params *p, float *out, int32_t row_begin,
                  int32_t row_end, int32_t threads) {
    if (!scene_ok(scene) || !p || !out) return 1;
    SceneData S(scene);
    if (threads > 0) omp_set_num_threads(threads);
    row_begin = std::max(0, row_begin);
    row_end = std::min(p->height, row_end);
#pragma omp parallel for schedule(guided)
    for (int py = row_begin; py < row_end; py++)
        for (int px = 0; px < p->width; px++) {
            Rng rng(p->seed, uint64_t(py) * uint64_t(p->width) + uint64_t(px));
            V3 c = S.render_pixel(*p, px, py, rng, nullptr);
            size_t k = (size_t(py) * size_t(p->width) + size_t(px)) * 3;
            out[k + 0] = c.x;
            out[k + 1] = c.y;
            out[k + 2] = c.z;
        }
    return 0;
}

int oracle_closest_hit(const lambda_scene_view *scene, const float *rays, int32_t n, lambda_hit *hits) {
    if (!scene_ok(scene)) return 1;
    SceneData S(scene);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        const float *q = rays + 6 * size_t(i);
        Ray r{mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]), spec(550)};
        Hit h;
        bool ok = S.hit(r, h, -1, -2, nullptr);
        lambda_hit &o = hits[i];
        o.hit = ok ? 1 : 0;
        o.instance = h.instance;
        o.object = h.object;
        o.t = h.t;
        o.point[0] = h.point.x; o.point[1] = h.point.y; o.point[2] = h.point.z;
        o.normal[0] = h.normal.x; o.normal[1] = h.normal.y; o.normal[2] = h.normal.z;
        o.u = h.u;
        o.v = h.v;
        o.front_facing = h.front ? 1 : 0;
    }
    return 0;
}

int oracle_occluded(const lambda_scene_view *scene, const float *rays, const int32_t *li, const int32_t *lo,
                    int32_t n, int32_t *occluded) {
    if (!scene_ok(scene)) return 1;
    SceneData S(scene);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        const float *q = rays + 6 * size_t(i);
        Ray r{mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]), spec(550)};
        Hit h;
        occluded[i] = S.hit(r, h, li[i], lo[i], nullptr) ? 1 : 0;
    }
    return 0;
}

int oracle_thinfilm(const lambda_scene_view *scene, int32_t material, const float *q, int32_t n, float *out_ref,
                    double *out_fp64, int32_t *out_ambiguous) {
    if (!scene || material < 0 || material >= scene->num_materials ||
        scene->material_type[material] != LAMBDA_MATERIAL_THINFILM)
        return 1;
    SceneData S(scene);
    for (int i = 0; i < n; i++) {
        const float *x = q + 7 * size_t(i);
        float c = x[0];
        float s = sqrtf(std::max(0.0f, 1 - c * c));
        Hit h;
        h.normal = mk(0, 0, 1);
        h.point = mk(x[3], x[4], x[5]);
        h.front = x[2] != 0;
        Ray r{mk(0, 0, 0), mk(s, 0, -c), spec(550)};
        double d = 0;
        bool amb = false;
        float f = S.thinfilm_reflectance(material, h, r, x[1], &d, &amb);
        if (out_ref) out_ref[i] = f;
        if (out_fp64) out_fp64[i] = d;
        if (out_ambiguous) out_ambiguous[i] = amb ? 1 : 0;
    }
    return 0;
}

void oracle_pcg32(uint64_t seed, uint64_t stream, int32_t n, uint32_t *out) {
    Rng r(seed, stream);
    for (int i = 0; i < n; i++) out[i] = r.step();
}

int oracle_path_stats(const lambda_scene_view *scene, const lambda_render_params *p, int64_t pixel_begin,
                      int64_t pixel_end, int64_t pixel_stride, int32_t threads, int64_t *counts) {
    if (!scene_ok(scene) || !p || !counts || pixel_stride < 1) return 1;
    SceneData S(scene);
    if (threads > 0) omp_set_num_threads(threads);
    int64_t total[13] = {0};
    const int64_t n = (pixel_end - pixel_begin + pixel_stride - 1) / pixel_stride;
#pragma omp parallel
    {
        Counters local;
#pragma omp for schedule(guided)
        for (int64_t i = 0; i < n; i++) {
            int64_t k = pixel_begin + i * pixel_stride;
            int px = int(k % p->width), py = int(k / p->width);
            Rng rng(p->seed, uint64_t(k));
            S.render_pixel(*p, px, py, rng, &local);
        }
#pragma omp critical
        for (int i = 0; i < 13; i++) total[i] += local.c[i];
    }
    for (int i = 0; i < 13; i++) counts[i] = total[i];
    return 0;
}

}  // extern "C"
