// Geometry helpers, BVH builder and light list of the host front end.
#include "scene_model.h"

#include <algorithm>
#include <map>
#include <numeric>

namespace lambda_host {

// ---------------------------------------------------------------- objects
// Constructors and bounds: scene/object.h:18-66,166-199.

Object Object::sphere(int m, const Vec3 &c, float r) {
    Object o;
    o.type = LAMBDA_OBJECT_SPHERE;
    o.material = m;
    o.a = c;
    o.radius = r;
    return o;
}

Object Object::quad(int m, const Vec3 &c, const Vec3 &h, const Vec3 &v) {
    Object o;
    o.type = LAMBDA_OBJECT_QUAD;
    o.material = m;
    o.a = c;
    o.h = h;
    o.v = v;
    o.n = cross(h, v);
    return o;
}

Object Object::tri(int m, const Vec3 &c, const Vec3 &h, const Vec3 &v, const Vec3 &n0,
                   const Vec3 &n1, const Vec3 &n2, float u0, float u1, float u2, float v0, float v1,
                   float v2) {
    Object o;
    o.type = LAMBDA_OBJECT_TRI;
    o.material = m;
    o.a = c;
    o.h = h;
    o.v = v;
    o.n = cross(h, v);
    // Only the first vertex normal decides whether shading normals were given.
    if (length(n0) < kEpsilon) {
        o.n0 = normalize(o.n);
        o.n1 = o.n0;
        o.n2 = o.n0;
    } else {
        o.n0 = n0;
        o.n1 = n1;
        o.n2 = n2;
    }
    o.u0 = u0; o.u1 = u1; o.u2 = u2;
    o.v0 = v0; o.v1 = v1; o.v2 = v2;
    return o;
}

Vec3 Object::bmin() const {
    const Vec3 eps(kEpsilon, kEpsilon, kEpsilon);
    switch (type) {
        case LAMBDA_OBJECT_SPHERE: return a - Vec3(radius, radius, radius);
        case LAMBDA_OBJECT_QUAD: return vmin(a, vmin(a + h, vmin(a + v, a + h + v))) - eps;
        default: return vmin(a, vmin(a + h, a + v)) - eps;
    }
}

Vec3 Object::bmax() const {
    const Vec3 eps(kEpsilon, kEpsilon, kEpsilon);
    switch (type) {
        case LAMBDA_OBJECT_SPHERE: return a + Vec3(radius, radius, radius);
        case LAMBDA_OBJECT_QUAD: return vmax(a, vmax(a + h, vmax(a + v, a + h + v))) + eps;
        default: return vmax(a, vmax(a + h, a + v)) + eps;
    }
}

Vec3 Object::center() const {
    switch (type) {
        case LAMBDA_OBJECT_SPHERE: return a;
        case LAMBDA_OBJECT_QUAD: return a + 0.5f * h + 0.5f * v;
        default: return a + (h + v) / 3.0f;
    }
}

float Object::bradius() const {
    switch (type) {
        case LAMBDA_OBJECT_SPHERE: return radius;
        case LAMBDA_OBJECT_QUAD: {
            float d1 = length_squared(h + v);
            float d2 = length_squared(h - v);
            return 0.5f * std::sqrt(std::fmax(d1, d2));
        }
        default: {
            Vec3 c = center();
            float d1 = length_squared(a - c);
            float d2 = length_squared(a + h - c);
            float d3 = length_squared(a + v - c);
            return std::sqrt(std::fmax(d1, std::fmax(d2, d3)));
        }
    }
}

float Object::area() const {
    switch (type) {
        case LAMBDA_OBJECT_SPHERE: return 4 * kPi * radius * radius;
        case LAMBDA_OBJECT_QUAD: return length(n);
        default: return length(n) / 2;
    }
}

// -------------------------------------------------------------- instances
// Object-to-world matrix T * Rz * Ry * Rx * S and its inverse
// (scene/instance.h:12-28).

Instance::Instance(int o, int c, int m, const Vec3 &t, const Vec3 &r, const Vec3 &s)
    : object(o), count(c), material(m) {
    Mat4 T = mat4(1, 0, 0, t.x, 0, 1, 0, t.y, 0, 0, 1, t.z, 0, 0, 0, 1);
    float sx = std::sin(r.x), cx = std::cos(r.x);
    float sy = std::sin(r.y), cy = std::cos(r.y);
    float sz = std::sin(r.z), cz = std::cos(r.z);
    Mat4 Rx = mat4(1, 0, 0, 0, 0, cx, -sx, 0, 0, sx, cx, 0, 0, 0, 0, 1);
    Mat4 Ry = mat4(cy, 0, sy, 0, 0, 1, 0, 0, -sy, 0, cy, 0, 0, 0, 0, 1);
    Mat4 Rz = mat4(cz, -sz, 0, 0, sz, cz, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1);
    Mat4 S = mat4(s.x, 0, 0, 0, 0, s.y, 0, 0, 0, 0, s.z, 0, 0, 0, 0, 1);
    M = T * Rz * Ry * Rx * S;
    Minv = inverse(M);
}

// World-space box of every object's 8 box corners (scene/instance.h:54-101).
Vec3 Instance::bmin(const std::vector<Object> &objects) const {
    Vec3 r(kMax, kMax, kMax);
    for (int i = 0; i < count; i++) {
        Vec3 p = objects[object + i].bmin(), q = objects[object + i].bmax();
        r = vmin(r, to_world(p));
        r = vmin(r, to_world(q));
        r = vmin(r, to_world(Vec3(p.x, p.y, q.z)));
        r = vmin(r, to_world(Vec3(p.x, q.y, p.z)));
        r = vmin(r, to_world(Vec3(p.x, q.y, q.z)));
        r = vmin(r, to_world(Vec3(q.x, p.y, p.z)));
        r = vmin(r, to_world(Vec3(q.x, p.y, q.z)));
        r = vmin(r, to_world(Vec3(q.x, q.y, p.z)));
    }
    return r;
}

Vec3 Instance::bmax(const std::vector<Object> &objects) const {
    Vec3 r(-kMax, -kMax, -kMax);
    for (int i = 0; i < count; i++) {
        Vec3 p = objects[object + i].bmin(), q = objects[object + i].bmax();
        r = vmax(r, to_world(p));
        r = vmax(r, to_world(q));
        r = vmax(r, to_world(Vec3(p.x, p.y, q.z)));
        r = vmax(r, to_world(Vec3(p.x, q.y, p.z)));
        r = vmax(r, to_world(Vec3(p.x, q.y, q.z)));
        r = vmax(r, to_world(Vec3(q.x, p.y, p.z)));
        r = vmax(r, to_world(Vec3(q.x, p.y, q.z)));
        r = vmax(r, to_world(Vec3(q.x, q.y, p.z)));
    }
    return r;
}

// bmin() and bmax() in one pass over the corners: the running minimum and
// maximum see the same values in the same order as the separate loops.
void Instance::bounds(const std::vector<Object> &objects, Vec3 &lo, Vec3 &hi) const {
    Vec3 a(kMax, kMax, kMax), b(-kMax, -kMax, -kMax);
    for (int i = 0; i < count; i++) {
        Vec3 p = objects[object + i].bmin(), q = objects[object + i].bmax();
        const Vec3 corners[8] = {p, q, Vec3(p.x, p.y, q.z), Vec3(p.x, q.y, p.z), Vec3(p.x, q.y, q.z),
                                 Vec3(q.x, p.y, p.z), Vec3(q.x, p.y, q.z), Vec3(q.x, q.y, p.z)};
        for (const Vec3 &c : corners) {
            Vec3 w = to_world(c);
            a = vmin(a, w);
            b = vmax(b, w);
        }
    }
    lo = a;
    hi = b;
}

Vec3 Instance::center(const std::vector<Object> &objects) const {
    Vec3 c(0, 0, 0);
    for (int i = 0; i < count; i++) c = c + to_world(objects[object + i].center());
    return c / float(count);
}

float Instance::radius(const std::vector<Object> &objects, int i) const {
    return objects[object + i].bradius() * length(Vec3(M(0, 0), M(1, 0), M(2, 0)));
}

float Instance::area(const std::vector<Object> &objects, int i) const {
    return objects[object + i].area() * length_squared(Vec3(M(0, 0), M(1, 0), M(2, 0)));
}

// --------------------------------------------------------------- spectra

float spectrum_average(const Spectrum &s) {
    float sum = 0;
    for (int i = 0; i < kBins; i++) sum += s[i];
    return sum / kBins;
}

float spectrum_min(const Spectrum &s) {
    float r = s[0];
    for (int i = 1; i < kBins; i++)
        if (s[i] < r) r = s[i];
    return r;
}

float spectrum_max(const Spectrum &s) {
    float r = s[0];
    for (int i = 1; i < kBins; i++)
        if (s[i] > r) r = s[i];
    return r;
}

// scene/texture.h:154-176,185-187.
float spectrum_texture_min(const Scene &s, int t) {
    const SpectrumTexture &x = s.spectrum_textures[t];
    switch (x.type) {
        case LAMBDA_SPECTRUM_CONSTANT: return spectrum_min(s.spectra[x.a]);
        case LAMBDA_SPECTRUM_CHECKER: return std::fmin(spectrum_min(s.spectra[x.a]), spectrum_min(s.spectra[x.b]));
        default: return s.scalar_textures[x.a].min();
    }
}

float spectrum_texture_max(const Scene &s, int t) {
    const SpectrumTexture &x = s.spectrum_textures[t];
    switch (x.type) {
        case LAMBDA_SPECTRUM_CONSTANT: return spectrum_max(s.spectra[x.a]);
        case LAMBDA_SPECTRUM_CHECKER: return std::fmax(spectrum_max(s.spectra[x.a]), spectrum_max(s.spectra[x.b]));
        default: return s.scalar_textures[x.a].max();
    }
}

float spectrum_texture_average(const Scene &s, int t) {
    const SpectrumTexture &x = s.spectrum_textures[t];
    switch (x.type) {
        case LAMBDA_SPECTRUM_CONSTANT: return spectrum_average(s.spectra[x.a]);
        case LAMBDA_SPECTRUM_CHECKER:
            return 0.5f * (spectrum_average(s.spectra[x.a]) + spectrum_average(s.spectra[x.b]));
        default: return s.scalar_textures[x.a].average();
    }
}

// -------------------------------------------------------------------- BVH
//
// Binned SAH over primitive boxes and centroids (scene/bvh.h:108-244).  The
// primitive bounds are evaluated once per primitive and reused: they are pure
// functions of the primitive, so memoising them changes no value, only the
// cost (the instance bounds loop over every object of the instance).

namespace {

struct PrimBounds {
    std::vector<Vec3> lo, hi, c;
};

int build_range(const PrimBounds &pb, std::vector<int> &idx, int start, int end,
                std::vector<Node> &nodes, int depth) {
    Vec3 bmin(kMax, kMax, kMax), bmax(-kMax, -kMax, -kMax);
    Vec3 cmin(kMax, kMax, kMax), cmax(-kMax, -kMax, -kMax);
    for (int i = start; i < end; i++) {
        int k = idx[i];
        bmin = vmin(bmin, pb.lo[k]);
        bmax = vmax(bmax, pb.hi[k]);
        cmin = vmin(cmin, pb.c[k]);
        cmax = vmax(cmax, pb.c[k]);
    }

    float best_cost = kMax;
    int best_axis = 0, best_split = -1;

    for (int axis = 0; axis < 3; axis++) {
        if (!(cmax[axis] - cmin[axis] >= kEpsilon)) continue;

        Vec3 bin_lo[kBinCount], bin_hi[kBinCount];
        int bin_n[kBinCount];
        for (int j = 0; j < kBinCount; j++) {
            bin_lo[j] = Vec3(kMax, kMax, kMax);
            bin_hi[j] = Vec3(-kMax, -kMax, -kMax);
            bin_n[j] = 0;
        }
        for (int j = start; j < end; j++) {
            int k = idx[j];
            int b = int(kBinCount * (pb.c[k][axis] - cmin[axis]) / (cmax[axis] - cmin[axis]));
            b = std::clamp(b, 0, kBinCount - 1);
            bin_lo[b] = vmin(bin_lo[b], pb.lo[k]);
            bin_hi[b] = vmax(bin_hi[b], pb.hi[k]);
            bin_n[b]++;
        }

        // Prefix (left) and suffix (right) sweeps over the bins.
        Vec3 l_lo[kBinCount], l_hi[kBinCount], r_lo[kBinCount], r_hi[kBinCount];
        int l_n[kBinCount], r_n[kBinCount];
        Vec3 run_lo(kMax, kMax, kMax), run_hi(-kMax, -kMax, -kMax);
        int run_n = 0;
        for (int j = 0; j < kBinCount; j++) {
            run_lo = vmin(run_lo, bin_lo[j]);
            run_hi = vmax(run_hi, bin_hi[j]);
            run_n += bin_n[j];
            l_lo[j] = run_lo; l_hi[j] = run_hi; l_n[j] = run_n;
        }
        run_lo = Vec3(kMax, kMax, kMax);
        run_hi = Vec3(-kMax, -kMax, -kMax);
        run_n = 0;
        for (int j = kBinCount - 1; j >= 0; j--) {
            run_lo = vmin(run_lo, bin_lo[j]);
            run_hi = vmax(run_hi, bin_hi[j]);
            run_n += bin_n[j];
            r_lo[j] = run_lo; r_hi[j] = run_hi; r_n[j] = run_n;
        }

        for (int j = 0; j < kBinCount - 1; j++) {
            if (l_n[j] == 0 || r_n[j + 1] == 0) continue;
            Vec3 le = l_hi[j] - l_lo[j];
            Vec3 re = r_hi[j + 1] - r_lo[j + 1];
            Vec3 te = bmax - bmin;
            float la = 2 * (le.x * le.y + le.x * le.z + le.y * le.z);
            float ra = 2 * (re.x * re.y + re.x * re.z + re.y * re.z);
            float ta = 2 * (te.x * te.y + te.x * te.z + te.y * te.z);
            float cost = la / ta * float(l_n[j]) + ra / ta * float(r_n[j + 1]);
            if (cost < best_cost) {
                best_cost = cost;
                best_axis = axis;
                best_split = j;
            }
        }
    }

    if (end - start == 1 || depth >= kMaxDepth || best_cost >= float(end - start) || best_split == -1) {
        Node leaf;
        leaf.min = bmin;
        leaf.max = bmax;
        leaf.index = start;
        leaf.count = end - start;
        nodes.push_back(leaf);
        return int(nodes.size()) - 1;
    }

    float plane = cmin[best_axis] + float(best_split + 1) * (cmax[best_axis] - cmin[best_axis]) / float(kBinCount);
    // std::partition (not stable_partition): the resulting order of the
    // indices is part of the output and matches the reference's libstdc++ call.
    auto it = std::partition(idx.begin() + start, idx.begin() + end,
                             [&](int k) { return pb.c[k][best_axis] < plane; });
    int mid = int(it - idx.begin());
    if (mid == start || mid == end) mid = start + (end - start) / 2;

    int current = int(nodes.size());
    Node inner;
    inner.min = bmin;
    inner.max = bmax;
    inner.index = -1;
    inner.count = 0;
    nodes.push_back(inner);
    build_range(pb, idx, start, mid, nodes, depth + 1);
    int right = build_range(pb, idx, mid, end, nodes, depth + 1);
    nodes[current].index = right;
    return current;
}

}  // namespace

void build_bvh(Scene &s, const std::vector<int> &offsets) {
    s.nodes.clear();
    if (s.objects.empty()) return;

    // Instance level.  Each instance's bounds loop over all its objects'
    // transformed box corners; instances are independent, so they are spread
    // over threads (each instance's own reduction keeps its sequential order).
    {
        PrimBounds pb;
        size_t n = s.instances.size();
        pb.lo.resize(n); pb.hi.resize(n); pb.c.resize(n);
#pragma omp parallel for schedule(dynamic, 1)
        for (long i = 0; i < long(n); i++) {
            s.instances[size_t(i)].bounds(s.objects, pb.lo[size_t(i)], pb.hi[size_t(i)]);
            pb.c[size_t(i)] = s.instances[size_t(i)].center(s.objects);
        }
        std::vector<int> idx(n);
        std::iota(idx.begin(), idx.end(), 0);
        if (n > 0) build_range(pb, idx, 0, int(n), s.nodes, 1);
        std::vector<Instance> ordered;
        ordered.reserve(n);
        for (size_t i = 0; i < n; i++) ordered.push_back(s.instances[idx[i]]);
        s.instances = std::move(ordered);
    }

    // One object-level BVH per object group, appended after the instance level.
    PrimBounds pb;
    size_t n = s.objects.size();
    pb.lo.resize(n); pb.hi.resize(n); pb.c.resize(n);
    for (size_t i = 0; i < n; i++) {
        pb.lo[i] = s.objects[i].bmin();
        pb.hi[i] = s.objects[i].bmax();
        pb.c[i] = s.objects[i].center();
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::map<int, int> root_of_offset;
    for (size_t g = 0; g + 1 < offsets.size(); g++) {
        root_of_offset[offsets[g]] = int(s.nodes.size());
        build_range(pb, idx, offsets[g], offsets[g + 1], s.nodes, 1);
    }
    for (Instance &inst : s.instances) inst.node = root_of_offset.at(inst.object);

    std::vector<Object> ordered;
    ordered.reserve(n);
    for (size_t i = 0; i < n; i++) ordered.push_back(s.objects[idx[i]]);
    s.objects = std::move(ordered);
}

// ------------------------------------------------------------------ lights

void build_lights(Scene &s) {
    // Incremental bounding sphere over every instanced object.
    Vec3 centre;
    float radius = 0;
    for (int i = 0; i < int(s.instances.size()); i++) {
        const Instance &inst = s.instances[i];
        for (int j = 0; j < inst.count; j++) {
            int object_index = inst.object + j;
            Vec3 oc = inst.center(s.objects, j);
            float orad = inst.radius(s.objects, j);
            Vec3 to = oc - centre;
            float dist = length(to);
            if (dist + radius <= orad) {
                centre = oc;
                radius = orad;
            } else if (dist + orad > radius) {
                float nr = 0.5f * (dist + radius + orad);
                centre = centre + (nr - radius) * to / dist;
                radius = nr;
            }
            int m = inst.material_of(s.objects, j);
            if (s.materials[m].emissive()) {
                s.light_instances.push_back(i);
                s.light_objects.push_back(object_index);
                float p = inst.area(s.objects, j) * spectrum_texture_average(s, s.materials[m].p[0]);
                s.light_powers.push_back(p);
                s.total_light_power += p;
            }
        }
    }
    // Environment: area of the bounding sphere times the mean radiance.
    s.light_instances.push_back(-1);
    s.light_objects.push_back(-1);
    float env = (4 * kPi * radius * radius) * spectrum_texture_average(s, s.background_texture);
    s.light_powers.push_back(env);
    s.total_light_power += env;
}

}  // namespace lambda_host
