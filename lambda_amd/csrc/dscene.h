// Device scene: layout in HBM, ray/primitive tests, two-level BVH traversal,
// textures, BSDFs, thin-film transfer matrices and light sampling.
//
// Layout:
//   wide[8 n .. 8 n + 7]: a node of the 4-wide BVH, 128 bytes, one cache line:
//     {lo.x[4]}, {hi.x[4]}, {lo.y[4]}, {hi.y[4]}, {lo.z[4]}, {hi.z[4]},
//     {bits(child)[4]}, {bits(count)[4]}; per slot count == 0: inner child at
//     wide node `child`, count > 0: leaf over [child, child + count) (objects,
//     or the first instance at the instance level), count < 0: empty slot.
//     Every box is a node box of the reference's own BVH (scene/bvh.h:14-28).
//   inst_root[i]: wide node of instance i's object-level BVH.
//   obj[4 n .. 4 n + 3] = {corner | centre, radius}, {horizontal, bits(type)},
//     {vertical, 0}, {cross(horizontal, vertical), 0}: one 64-byte record, so
//     a triangle test reads one cache line; obj_s[n][4] is the triangle
//     shading record, read once per closest hit.
//   inst[n][8] = {M rows 0-3, M^-1 rows 0-3}, inst_i[n] = {object, count,
//     BLAS root, material}
//   spectra are 471-bin float tables, complex spectra 471-bin float2 tables.
#pragma once

#include "dmath.h"
#include "ffloat.h"
#include "lambda_hip.h"

namespace lam {

struct DeviceScene {
    const float4 *wide;     // 4-wide BVH nodes, 8 float4 each
    const int *inst_root;   // [instances] wide root of the object-level BVH
    const float4 *obj;      // [objects][4]
    const float4 *obj_s;    // [objects][4] triangle shading records
    const int *obj_mat;
    const float4 *inst;
    const int4 *inst_i;
    const int4 *mat;        // {type, p0, p1, p2}
    const float4 *stex;     // {bits(type), p0, p1, p2}
    const int4 *xtex;       // {type, a, b, bits(scale)}
    const float *spectra;   // [S][471]
    const float2 *cspectra; // [C][471]
    const int *props;
    const int2 *light_io;   // {instance, object}
    const float *light_pow;
    const float *cie;       // [471][3]
    const int *perm;        // [256]
    const float4 *worley_points;  // feature points of all tables
    const int4 *worley_box;       // [scalar textures] {x0, y0, z0, offset}
    const int4 *worley_dim;       // [scalar textures] {nx, ny, nz, 0}; nx = 0: no table
    const int *tlas_rank;   // [instances] order of the instance's leaf in the reference's DFS
    const int *obj_rank;    // [objects] order within its object-level BVH in that DFS
    const int4 *mat_tex;    // [materials] xtex record of the material's texture (Lambertian, mirror, emissive)
    const float *tex_avg;   // [spectrum textures] mean over wavelength (spectrum_texture_average)
    float4 cam_pos, cam_corner, cam_h, cam_v;
    float total_power;
    int num_lights, bg_tex;
    int num_nodes, num_objects, num_instances;
};

// ------------------------------------------------------------------ records

struct Hit {
    int inst, obj;
    float t;
    V3 p, n;
    float u, v;
    bool front;
};

// Pins a loaded value's load where this is called: a load whose value is
// used only inside a branch is moved into that branch by the compiler, where
// it costs a memory round trip of its own.  Calling this after a group of
// loads keeps them in flight together (one round trip for the group).
LAM_FN void keep(float a, float b = 0, float c = 0, float d = 0) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" ::"v"(a), "v"(b), "v"(c), "v"(d));
#else
    (void)a, (void)b, (void)c, (void)d;
#endif
}
LAM_FN void keep(int a) { keep(as_float(a)); }
LAM_FN void keep(float4 v) { keep(v.x, v.y, v.z, v.w); }

LAM_FN float dense(const float *s, float lambda) {
    // 1 nm table, linear interpolation (math/spectrum.h:170-188) in float.
    float index = lambda - 360;
    int lo = int(index);
    if (lo >= kBins - 1) return s[kBins - 1];
    float a = s[lo], b = s[lo + 1];
    return a + (index - float(lo)) * (b - a);
}

LAM_FN float2 dense_c(const float2 *s, float lambda) {
    float index = lambda - 360;
    int lo = int(index);
    if (lo >= kBins - 1) return s[kBins - 1];
    float2 a = s[lo], b = s[lo + 1];
    float w = index - float(lo);
    return make_float2(a.x + w * (b.x - a.x), a.y + w * (b.y - a.y));
}

// ---------------------------------------------------------------- geometry
// Ray/primitive tests (scene/object.h:301-400).  `best` is the current
// closest distance; a hit must lie in [1e-4, best).  These return the
// candidate distance only; hit_record() rebuilds the full record for the
// winning primitive with the same arithmetic.

LAM_FN bool test_sphere(float4 a, V3 o, V3 d, float best, float &t_out) {
    V3 oc = o - v3(a);
    float A = len2(d);
    float B = 2 * dot(d, oc);
    float C = dot(oc, oc) - a.w * a.w;
    float D = B * B - 4 * A * C;
    if (D < 0) return false;
    D = sqrtf(D);
    float t1 = (-B - D) / (2 * A);
    float t2 = (-B + D) / (2 * A);
    if (t1 < kEps && t2 < kEps) return false;
    float t = (t1 < kEps) ? t2 : t1;
    if (t >= best) return false;
    t_out = t;
    return true;
}

LAM_FN bool test_quad(V3 c, V3 h, V3 v, V3 nrm, V3 o, V3 d, float best, float &t_out,
                                          float &alpha, float &beta) {
    V3 n = normalize(nrm);
    float den = dot(n, d);
    if (fabsf(den) < kEps) return false;
    float t = dot(n, c - o) / den;
    if (t < kEps || t >= best) return false;
    V3 p = o + t * d;
    V3 proj = p - c;
    alpha = dot(nrm, cross(proj, v)) / len2(nrm);
    beta = dot(nrm, cross(h, proj)) / len2(nrm);
    if (alpha < 0 || alpha > 1 || beta < 0 || beta > 1) return false;
    t_out = t;
    return true;
}

LAM_FN bool test_tri(V3 c, V3 h, V3 v, V3 o, V3 d, float best, float &t_out, float &alpha,
                                         float &beta) {
    V3 rcv = cross(d, v);
    float det = dot(h, rcv);
    if (fabsf(det) < kEps2) return false;
    float inv = 1 / det;
    V3 oc = o - c;
    alpha = inv * dot(oc, rcv);
    if (alpha < 0 || alpha > 1) return false;
    V3 och = cross(oc, h);
    beta = inv * dot(d, och);
    if (beta < 0 || alpha + beta > 1) return false;
    float t = inv * dot(v, och);
    if (t < kEps || t >= best) return false;
    t_out = t;
    return true;
}

// object_record from the object's records already loaded: R = obj[4k ..
// 4k + 3], SH = its shading record.
LAM_FN bool object_record_rec(const float4 (&R)[4], const float4 (&SH)[4], int k, V3 o, V3 d, float best, Hit &h) {
    const float4 a = R[0], b = R[1];
    int type = as_int(b.w);
    float t, al = 0, be = 0;
    if (type == LAMBDA_OBJECT_SPHERE) {
        if (!test_sphere(a, o, d, best, t)) return false;
        h.t = t;
        h.p = o + t * d;
        h.n = normalize(h.p - v3(a));
        h.u = (atan2f(-h.n.z, h.n.x) + kPi) / (2 * kPi);
        h.v = acosf(-h.n.y) / kPi;
        h.front = dot(d, h.n) < 0;
    } else {
        const float4 c = R[2], nn = R[3];
        V3 nrm = v3(nn);
        if (type == LAMBDA_OBJECT_TRI) {
            if (!test_tri(v3(a), v3(b), v3(c), o, d, best, t, al, be)) return false;
            const float4 s0 = SH[0], s1 = SH[1], s2 = SH[2], s3 = SH[3];
            V3 n0 = v3(s0.x, s0.y, s0.z), n1 = v3(s0.w, s1.x, s1.y), n2 = v3(s1.z, s1.w, s2.x);
            float w = 1 - al - be;
            h.t = t;
            h.p = o + t * d;
            h.n = normalize(w * n0 + al * n1 + be * n2);
            h.u = w * s2.y + al * s2.z + be * s2.w;
            h.v = w * s3.x + al * s3.y + be * s3.z;
            h.front = dot(d, nrm) < 0;
        } else {
            if (!test_quad(v3(a), v3(b), v3(c), nrm, o, d, best, t, al, be)) return false;
            h.t = t;
            h.p = o + t * d;
            h.n = normalize(nrm);
            h.u = al;
            h.v = be;
            h.front = dot(d, h.n) < 0;
        }
    }
    if (!h.front) h.n = h.n * -1;
    h.obj = k;
    return true;
}

LAM_FN void load_object(const DeviceScene &S, int k, float4 (&R)[4], float4 (&SH)[4]) {
    const float4 *r = S.obj + 4 * size_t(k), *sh = S.obj_s + 4 * size_t(k);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        R[i] = r[i];
        SH[i] = sh[i];
    }
}

// Full intersection record of object k for a local-space ray, as the
// reference's intersect* fills it (scene/object.h:301-400).
LAM_FN bool object_record(const DeviceScene &S, int k, V3 o, V3 d, float best, Hit &h) {
    float4 R[4], SH[4];
    load_object(S, k, R, SH);
    return object_record_rec(R, SH, k, o, d, best, h);
}

LAM_FN void local_ray(const DeviceScene &S, int inst, V3 o, V3 d, V3 &lo, V3 &ld) {
    const float4 *m = S.inst + 8 * size_t(inst);
    lo = xform(m + 4, o, 1);
    ld = xform(m + 4, d, 0);
}

// World-space record of (instance, object) for a world ray: local test, then
// point by M and normal by the transposed inverse (scene/scene.h:101-104).
// The transform and the object's intersection and shading records are read
// before the test, so that none of their loads waits on the test's branches.
LAM_FN bool hit_record(const DeviceScene &S, int inst, int obj, V3 o, V3 d, Hit &h) {
    const float4 *mp = S.inst + 8 * size_t(inst);
    float4 M[8];
#pragma unroll
    for (int i = 0; i < 8; i++) M[i] = mp[i];
    float4 R[4], SH[4];
    load_object(S, obj, R, SH);
    const V3 lo = xform(M + 4, o, 1), ld = xform(M + 4, d, 0);
    if (!object_record_rec(R, SH, obj, lo, ld, kMax, h)) return false;
    h.inst = inst;
    h.p = xform(M, h.p, 1);
    h.n = normalize(xform_t(M + 4, h.n, 0));
    return true;
}

// ------------------------------------------------------------- traversal
//
// Two-level BVH traversal (scene/scene.h:51-124) over a 4-wide BVH whose
// boxes and leaves are the reference's own: each wide node holds up to four
// descendants of one reference node, chosen at upload by expanding the
// child of largest surface area.  A reference node's box contains its
// descendants' boxes exactly, and the slab test is monotone in the box
// bounds, so a descendant whose box is hit (and not beyond the closest hit)
// has every intermediate box hit as well: testing the four descendants
// directly finds the same candidates as the reference's walk.
//
// At a wide node all four boxes are tested; object leaves are intersected
// at once, inner children (and instance leaves at the instance level) are
// sorted by entry distance, the nearest is entered and the others pushed
// with their entry distance, so that a popped entry beyond the closest hit
// is dropped without a memory access.  Only the first instance of an
// instance-level leaf is entered, as in the reference (scene/scene.h:79-86).
//
// Visiting order changes which of several primitives at exactly the same
// distance is met first.  The reference keeps the first one in its own
// depth-first order (right child before left, scene/scene.h:112-119), so on
// an exact tie the candidate with the smaller reference rank wins:
// rank = (instance-level leaf order << 32) | object order within its BVH,
// both precomputed on the host by replaying that order.

// Slab test of one box (scene/bvh.h:30-52): entry distance clamped at 0,
// exit clamped at 1e30.
LAM_FN bool slab(float lox, float loy, float loz, float hix, float hiy, float hiz, V3 o, V3 inv, float &tnear) {
    float t1 = 0, t2 = kMax;
    float a = (lox - o.x) * inv.x, b = (hix - o.x) * inv.x;
    if (inv.x < 0) { float s = a; a = b; b = s; }
    t1 = fmaxf(a, t1);
    t2 = fminf(b, t2);
    a = (loy - o.y) * inv.y;
    b = (hiy - o.y) * inv.y;
    if (inv.y < 0) { float s = a; a = b; b = s; }
    t1 = fmaxf(a, t1);
    t2 = fminf(b, t2);
    a = (loz - o.z) * inv.z;
    b = (hiz - o.z) * inv.z;
    if (inv.z < 0) { float s = a; a = b; b = s; }
    t1 = fmaxf(a, t1);
    t2 = fminf(b, t2);
    tnear = t1;
    return t1 <= t2 && t2 >= 0;
}

LAM_FN uint64_t ref_rank(const DeviceScene &S, int inst, int obj) {
    return (uint64_t(uint32_t(S.tlas_rank[inst])) << 32) | uint64_t(uint32_t(S.obj_rank[obj]));
}

// Test of object k, whose record starts {a, b, c}, against the (local) ray;
// any = true: occlusion mode (any hit within best counts), else closest-hit
// mode: a closer hit always wins; a hit at exactly the current distance wins
// when it comes first in the reference's order.
LAM_FN bool test_object_mode(const DeviceScene &S, bool any, int k, float4 a, float4 b, float4 c, V3 o, V3 d,
                             float &best, int inst, int &hit_inst, int &hit_obj) {
    int type = as_int(b.w);
    float limit = any ? best : nextafterf(best, INFINITY);
    float t, al, be;
    if (type == LAMBDA_OBJECT_SPHERE) {
        if (!test_sphere(a, o, d, limit, t)) return false;
    } else if (type == LAMBDA_OBJECT_TRI) {
        if (!test_tri(v3(a), v3(b), v3(c), o, d, limit, t, al, be)) return false;
    } else {
        float4 n = S.obj[4 * size_t(k) + 3];
        if (!test_quad(v3(a), v3(b), v3(c), v3(n), o, d, limit, t, al, be)) return false;
    }
    if (!any && t == best && hit_obj >= 0 && ref_rank(S, inst, k) > ref_rank(S, hit_inst, hit_obj))
        return false;
    best = t;
    hit_inst = inst;
    hit_obj = k;
    return true;
}

// Entries that wait for a visit: a wide node (code >= 0) or an instance
// (code = -1 - instance), with the entry distance of its box.  The top
// kRing entries of each lane live in LDS (entry j at slot j mod kRing,
// stride = block size, so a wave touches 64 consecutive words); deeper
// entries are moved to a global area laid out [entry][thread] when the ring
// is full and brought back when the stack shrinks to them.  The host sizes
// the global area from the exact worst-case depth of the tree.
//
// In device code the ring pointers carry the LDS address space: with plain
// (generic) pointers the compiler turns the pop's two candidate reads (ring
// or global area) into one flat load through a selected pointer, which is
// slower than an LDS read and waits on both memory counters.
#if defined(__HIP_DEVICE_COMPILE__)
#define LAM_LDS __attribute__((address_space(3)))
#else
#define LAM_LDS
#endif

template <int kRing>
struct TraversalStack {
    LAM_LDS int *lcode;     // LDS, kRing * stride
    LAM_LDS float *ldist;   // LDS, kRing * stride
    int stride;
    int2 *spill;         // global, entry j of this thread at spill[j * spill_stride]
    size_t spill_stride;
    int sp = 0;          // entries on the stack
    int lo = 0;          // entries below lo are in the global area

    LAM_FN void push(int code, float dist) {
        if (sp - lo == kRing) {
            const int s = (lo & (kRing - 1)) * stride;
            spill[size_t(lo) * spill_stride] = make_int2(lcode[s], as_int(ldist[s]));
            lo++;
        }
        const int s = (sp & (kRing - 1)) * stride;
        lcode[s] = code;
        ldist[s] = dist;
        sp++;
    }
    LAM_FN void pop(int &code, float &dist) {
        sp--;
        if (sp < lo) {
            int2 e = spill[size_t(sp) * spill_stride];
            code = e.x;
            dist = as_float(e.y);
            lo = sp;
        } else {
            const int s = (sp & (kRing - 1)) * stride;
            code = lcode[s];
            dist = ldist[s];
        }
    }
};

// Sorts up to four (distance, code) pairs by distance (a five-exchange
// network); misses carry +inf.
LAM_FN void sort4(float (&t)[4], int (&c)[4]) {
#define LAM_CSWAP(i, j)                                            \
    if (t[j] < t[i]) {                                             \
        float tt = t[i]; t[i] = t[j]; t[j] = tt;                   \
        int cc = c[i]; c[i] = c[j]; c[j] = cc;                     \
    }
    LAM_CSWAP(0, 1)
    LAM_CSWAP(2, 3)
    LAM_CSWAP(0, 2)
    LAM_CSWAP(1, 3)
    LAM_CSWAP(1, 2)
#undef LAM_CSWAP
}

// kAny = false: closest hit; returns true if anything was hit, with
//   (hit_inst, hit_obj, best).
// kAny = true: occlusion; best is the light distance, (skip_inst, skip_obj)
//   the light itself; returns true on the first blocking hit.
template <bool kAny, int kRing>
LAM_FN bool traverse(const DeviceScene &S, V3 o, V3 d, float &best, int &hit_inst, int &hit_obj, int skip_inst,
                     int skip_obj, TraversalStack<kRing> &stk);

// One lane's occlusion ray (when `occl`) and then its closest-hit ray (when
// `closest`) in one loop.  A lane moves on to its closest-hit ray as soon
// as its occlusion ray is resolved, so in a wave the lanes without an
// occlusion ray traverse for their closest hits while the others finish
// theirs, instead of the whole wave running a separate any-hit traversal
// for a few lanes.  Each ray gets exactly the answer of traverse<true> and
// traverse<false>: the visiting order, the entry-distance pruning and the
// ranked tie rule are the same, only the mode is chosen at run time.
//   occlusion ray: (so, sd), light distance s_dist, light (skip_inst,
//     skip_obj); blocked = true when something lies before the light.
//   closest-hit ray: origin and direction read from *co, *cd when the
//     occlusion ray is done (or at once); (hit_inst, hit_obj, best) as
//     traverse<false> returns them, hit_inst = hit_obj = -1 on a miss.
template <int kRing>
LAM_FN void traverse_pair(const DeviceScene &S, bool occl, V3 so, V3 sd, float s_dist, int skip_inst, int skip_obj,
                          bool &blocked, bool closest, const float4 *co, const float4 *cd, float &best,
                          int &hit_inst, int &hit_obj, TraversalStack<kRing> &stk);

// Resumable form of traverse_pair's loop for kernels that interleave
// traversal with other work: begin_*() sets a ray up, each step() visits one
// wide node (and tests the objects of its hit leaves) and returns true once
// the ray is resolved.  Same visiting order, pruning and tie rule as
// traverse<kAny>, so the results are the same.
struct RayWalk {
    V3 o, d, inv;     // world ray
    V3 ro, rd, ri;    // ray of the current level
    float lim;        // closest hit so far, or the light distance
    int hi, ho;       // closest hit so far (closest-hit mode)
    int inst, blas_base, cur;
    int skip_inst, skip_obj;
    bool any, in_blas, blocked;

    LAM_FN void begin(bool any_hit, V3 o_, V3 d_, float lim_, int skip_i, int skip_o) {
        any = any_hit;
        o = o_;
        d = d_;
        inv = v3(1 / d.x, 1 / d.y, 1 / d.z);
        ro = o;
        rd = d;
        ri = inv;
        lim = lim_;
        hi = ho = -1;
        inst = -1;
        blas_base = 0;
        cur = 0;
        skip_inst = skip_i;
        skip_obj = skip_o;
        in_blas = false;
        blocked = false;
    }

    template <int kRing>
    LAM_FN bool step(const DeviceScene &S, TraversalStack<kRing> &stk) {
        if (cur < 0) {
            inst = -1 - cur;
            local_ray(S, inst, o, d, ro, rd);
            ri = v3(1 / rd.x, 1 / rd.y, 1 / rd.z);
            in_blas = true;
            blas_base = stk.sp;
            cur = S.inst_root[inst];
        }
        const float4 *w = S.wide + 8 * size_t(cur);
        const float4 lx = w[0], hx = w[1], ly = w[2], hy = w[3], lz = w[4], hz = w[5];
        const float4 cw = w[6], nw = w[7];
        const float LX[4] = {lx.x, lx.y, lx.z, lx.w}, HX[4] = {hx.x, hx.y, hx.z, hx.w};
        const float LY[4] = {ly.x, ly.y, ly.z, ly.w}, HY[4] = {hy.x, hy.y, hy.z, hy.w};
        const float LZ[4] = {lz.x, lz.y, lz.z, lz.w}, HZ[4] = {hz.x, hz.y, hz.z, hz.w};
        const int CH[4] = {as_int(cw.x), as_int(cw.y), as_int(cw.z), as_int(cw.w)};
        const int CN[4] = {as_int(nw.x), as_int(nw.y), as_int(nw.z), as_int(nw.w)};
        float key[4];
        int code[4];
        unsigned leaves = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float tn;
            // all four boxes are tested, empty slots too (their result is
            // masked): with the test behind the slot check the compiler
            // issues the box loads inside the branch, a second round trip
            // to memory per node
            const bool in = slab(LX[j], LY[j], LZ[j], HX[j], HY[j], HZ[j], ro, ri, tn);
            bool hit = in & (CN[j] >= 0) & !(tn > lim);
            key[j] = INFINITY;
            code[j] = 0;
            if (!hit) continue;
            if (CN[j] == 0) {
                key[j] = tn;
                code[j] = CH[j];
            } else if (!in_blas) {
                key[j] = tn;
                code[j] = -1 - CH[j];
            } else {
                leaves |= 1u << j;
            }
        }
        int k = 0, end = 0;
        while (true) {
            if (k >= end) {
                if (leaves == 0) break;
                const int j = __builtin_ctz(leaves);
                leaves &= leaves - 1;
                k = j == 0 ? CH[0] : j == 1 ? CH[1] : j == 2 ? CH[2] : CH[3];
                end = k + (j == 0 ? CN[0] : j == 1 ? CN[1] : j == 2 ? CN[2] : CN[3]);
            }
            const bool two = k + 1 < end;
            const float4 *r0 = S.obj + 4 * size_t(k);
            const float4 *r1 = two ? r0 + 4 : r0;
            const float4 a0 = r0[0], b0 = r0[1], c0 = r0[2];
            const float4 a1 = r1[0], b1 = r1[1], c1 = r1[2];
            if (!(any && inst == skip_inst && k == skip_obj) &&
                test_object_mode(S, any, k, a0, b0, c0, ro, rd, lim, inst, hi, ho) && any) {
                blocked = true;
                return true;
            }
            if (two && !(any && inst == skip_inst && k + 1 == skip_obj) &&
                test_object_mode(S, any, k + 1, a1, b1, c1, ro, rd, lim, inst, hi, ho) && any) {
                blocked = true;
                return true;
            }
            k += 2;
        }
        sort4(key, code);
        int m = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) m += !(key[j] > lim) ? 1 : 0;
        if (m > 0) {
            for (int j = m - 1; j >= 1; j--) stk.push(code[j], key[j]);
            cur = code[0];
            return false;
        }
        while (true) {
            if (in_blas && stk.sp == blas_base) {
                in_blas = false;
                ro = o;
                rd = d;
                ri = inv;
                continue;
            }
            if (stk.sp == 0) return true;
            int c;
            float t;
            stk.pop(c, t);
            if (!(t > lim)) {
                cur = c;
                return false;
            }
        }
    }
};

template <int kRing>
LAM_FN void traverse_pair(const DeviceScene &S, bool occl, V3 so, V3 sd, float s_dist, int skip_inst, int skip_obj,
                          bool &blocked, bool closest, const float4 *co, const float4 *cd, float &best,
                          int &hit_inst, int &hit_obj, TraversalStack<kRing> &stk) {
    blocked = false;
    hit_inst = hit_obj = -1;
    best = kMax;
    if (!occl && !closest) return;
    RayWalk w;
    if (occl) w.begin(true, so, sd, s_dist, skip_inst, skip_obj);
    else w.begin(false, v3(*co), v3(*cd), kMax, -1, -2);
    stk.sp = stk.lo = 0;
    while (true) {
        if (!w.template step<kRing>(S, stk)) continue;
        if (w.any) {
            blocked = w.blocked;
            if (!closest) return;
            w.begin(false, v3(*co), v3(*cd), kMax, -1, -2);
            stk.sp = stk.lo = 0;
            continue;
        }
        hit_inst = w.hi;
        hit_obj = w.ho;
        best = w.lim;
        return;
    }
}

// traverse<kAny> through RayWalk: one ray, stepped until it is resolved.
template <bool kAny, int kRing>
LAM_FN bool traverse(const DeviceScene &S, V3 o, V3 d, float &best, int &hit_inst, int &hit_obj, int skip_inst,
                     int skip_obj, TraversalStack<kRing> &stk) {
    RayWalk w;
    w.begin(kAny, o, d, best, skip_inst, skip_obj);
    while (!w.template step<kRing>(S, stk)) {
    }
    if (kAny) return w.blocked;
    if (w.hi < 0) return false;
    best = w.lim;
    hit_inst = w.hi;
    hit_obj = w.ho;
    return true;
}

// --------------------------------------------------------------- textures

//
// The out-of-line functions take the tables they read as pointer arguments,
// never the scene by reference: a reference would make the kernel keep its
// scene record in scratch memory, and every table pointer read back from
// there loses its global address space, so the traversal's loads would be
// issued as flat loads (which also make every LDS stack access wait for
// them).

LAM_OUTLINE float scalar_texture_at(const float4 *stex, const int *perm, const float4 *worley_points,
                                    const int4 *worley_box, const int4 *worley_dim, int t, V3 p) {
    float4 x = stex[t];
    int type = as_int(x.x);
    if (type == LAMBDA_SCALAR_CONSTANT) return x.y;
    if (type == LAMBDA_SCALAR_PERLIN) return x.y + (x.z - x.y) * perlin(perm, p * x.w);
    const int4 box = worley_box[t], dim = worley_dim[t];
    WorleyTable tab{dim.x > 0 ? worley_points + box.w : nullptr, box.x, box.y, box.z, dim.x, dim.y, dim.z};
    return x.y + (x.z - x.y) * worley(p * x.w, tab);
}

LAM_FN float scalar_texture(const DeviceScene &S, int t, V3 p) {
    return scalar_texture_at(S.stex, S.perm, S.worley_points, S.worley_box, S.worley_dim, t, p);
}

// Constant or checker spectrum texture x at one wavelength
// (scene/texture.h:178-210); checker cells are in (u, v) units.
LAM_FN float spectrum_texture_tabulated(const DeviceScene &S, int4 x, const Hit &h, float lambda) {
    int s = x.y;
    if (x.x == LAMBDA_SPECTRUM_CHECKER) {
        float scale = as_float(x.w);
        int cx = int(floorf(h.u / scale));
        int cy = int(floorf(h.v / scale));
        s = (cx + cy) % 2 == 0 ? x.y : x.z;
    }
    return dense(S.spectra + size_t(s) * kBins, lambda);
}

// Spectrum texture at one wavelength (scene/texture.h:178-210).
LAM_FN float spectrum_texture(const DeviceScene &S, int t, const Hit &h, float lambda) {
    int4 x = S.xtex[t];
    if (x.x == LAMBDA_SPECTRUM_SCALAR) return scalar_texture(S, x.y, h.p);
    return spectrum_texture_tabulated(S, x, h, lambda);
}

// Texture that does not depend on the wavelength: evaluated once per vertex.
LAM_FN bool spectrum_texture_is_scalar(const DeviceScene &S, int t) {
    return S.xtex[t].x == LAMBDA_SPECTRUM_SCALAR;
}

// The four wavelengths of a vertex; a scalar texture is evaluated once (it
// does not depend on the wavelength), so only one copy of its code is
// inlined per call site.
// x: the texture's xtex record.
LAM_FN S4 spectrum_texture4_rec(const DeviceScene &S, int4 x, const Hit &h, const S4 &lam) {
    S4 r;
    if (x.x == LAMBDA_SPECTRUM_SCALAR) return s4(scalar_texture(S, x.y, h.p));
#pragma unroll
    for (int j = 0; j < kHero; j++) r.v[j] = spectrum_texture_tabulated(S, x, h, lam.v[j]);
    return r;
}

LAM_FN S4 spectrum_texture4(const DeviceScene &S, int t, const Hit &h, const S4 &lam) {
    return spectrum_texture4_rec(S, S.xtex[t], h, lam);
}

// Environment lookup by direction: equirectangular (u, v), point = direction
// (scene/background.h:81-95).
LAM_FN S4 background(const DeviceScene &S, V3 d, const S4 &lam) {
    Hit h;
    h.p = d;
    h.u = (atan2f(d.z, d.x) + kPi) / (2 * kPi);
    h.v = acosf(d.y) / kPi;
    return spectrum_texture4(S, S.bg_tex, h, lam);
}

LAM_FN float spectrum_average(const DeviceScene &S, int s) {
    const float *p = S.spectra + size_t(s) * kBins;
    float sum = 0;
    for (int i = 0; i < kBins; i++) sum += p[i];
    return sum / kBins;
}

// Mean emitted radiance of a spectrum texture (scene/texture.h:154-176).
LAM_FN float spectrum_texture_average(const DeviceScene &S, int t) {
    int4 x = S.xtex[t];
    if (x.x == LAMBDA_SPECTRUM_CONSTANT) return spectrum_average(S, x.y);
    if (x.x == LAMBDA_SPECTRUM_CHECKER) return float(0.5) * (spectrum_average(S, x.y) + spectrum_average(S, x.z));
    float4 st = S.stex[x.y];
    return as_int(st.x) == LAMBDA_SCALAR_CONSTANT ? st.y : (st.y + st.z) / 2;
}

// ---------------------------------------------------------------- objects
// Light sampling on objects (scene/object.h:201-299) and instances
// (scene/instance.h:103-111).

LAM_FN float object_area(const DeviceScene &S, int k) {
    float4 a = S.obj[4 * size_t(k)];
    int type = as_int(S.obj[4 * size_t(k) + 1].w);
    if (type == LAMBDA_OBJECT_SPHERE) return 4 * kPi * a.w * a.w;
    float l = len(v3(S.obj[4 * size_t(k) + 3]));
    return type == LAMBDA_OBJECT_QUAD ? l : l / 2;
}

LAM_FN float object_pdf(const DeviceScene &S, int k, V3 p, V3 d) {
    float4 a = S.obj[4 * size_t(k)];
    float4 b = S.obj[4 * size_t(k) + 1];
    int type = as_int(b.w);
    if (type == LAMBDA_OBJECT_SPHERE) {
        V3 oc = v3(a) - p;
        float dist2 = len2(oc);
        float r2 = a.w * a.w;
        if (dist2 <= r2) return 1 / (4 * kPi);
        float cmax = sqrtf(1 - r2 / dist2);
        float ct = dot(normalize(oc), d);
        if (ct < cmax) return 0;
        return 1 / (2 * kPi * (1 - cmax));
    }
    V3 c = v3(a), h = v3(b), v = v3(S.obj[4 * size_t(k) + 2]), n = v3(S.obj[4 * size_t(k) + 3]);
    if (type == LAMBDA_OBJECT_QUAD) {
        float ct = dot(normalize(n), d);
        if (fabsf(ct) < kEps) return 0;
        float t = dot(c - p, normalize(n)) / ct;
        if (t < kEps) return 0;
        V3 local = (p + d * t) - c;
        float u = dot(n, cross(local, v)) / len2(n);
        float w = dot(n, cross(h, local)) / len2(n);
        if (u < 0 || u > 1 || w < 0 || w > 1) return 0;
        return t * t / (fabsf(ct) * len(n));
    }
    V3 rcv = cross(d, v);
    float det = dot(h, rcv);
    if (fabsf(det) < kEps2) return 0;
    float inv = 1 / det;
    V3 oc = p - c;
    float al = inv * dot(oc, rcv);
    if (al < 0 || al > 1) return 0;
    V3 och = cross(oc, h);
    float be = inv * dot(d, och);
    if (be < 0 || al + be > 1) return 0;
    float t = inv * dot(v, och);
    if (t < kEps) return 0;
    float area = len(n) * 0.5f;
    float ct = dot(normalize(n), d);
    return t * t / (fabsf(ct) * area);
}

LAM_FN V3 object_sample(const DeviceScene &S, int k, V3 p, Pcg &rng) {
    float4 a = S.obj[4 * size_t(k)];
    float4 b = S.obj[4 * size_t(k) + 1];
    int type = as_int(b.w);
    if (type == LAMBDA_OBJECT_SPHERE) {
        V3 oc = v3(a) - p;
        float dist2 = len2(oc);
        float r2 = a.w * a.w;
        if (dist2 <= r2) return random_unit_vector(rng);
        float cmax = sqrtf(1 - r2 / dist2);
        return random_in_cone(oc, cmax, rng);
    }
    float u = rng.next();
    float v = rng.next();
    if (type == LAMBDA_OBJECT_TRI && u + v > 1) {
        u = 1 - u;
        v = 1 - v;
    }
    V3 s = v3(a) + v3(b) * u + v3(S.obj[4 * size_t(k) + 2]) * v;
    return normalize(s - p);
}

LAM_FN int material_of(const DeviceScene &S, int inst, int obj) {
    int m = S.inst_i[inst].w;
    return m >= 0 ? m : S.obj_mat[obj];
}

LAM_FN float instance_area(const DeviceScene &S, int inst, int obj) {
    const float4 *m = S.inst + 8 * size_t(inst);
    return object_area(S, obj) * len2(v3(m[0].x, m[1].x, m[2].x));
}

LAM_FN float instance_pdf(const DeviceScene &S, int inst, int obj, V3 p, V3 d) {
    const float4 *m = S.inst + 8 * size_t(inst);
    return object_pdf(S, obj, xform(m + 4, p, 1), normalize(xform(m + 4, d, 0)));
}

LAM_FN V3 instance_sample(const DeviceScene &S, int inst, int obj, V3 p, Pcg &rng) {
    const float4 *m = S.inst + 8 * size_t(inst);
    return normalize(xform(m, object_sample(S, obj, xform(m + 4, p, 1), rng), 0));
}

// --------------------------------------------------------------- thin film
//
// Reflectance of a layer stack.
//
// Lambda forms R = (|M01/M00|^2_s + |M01/M00|^2_p) / 2 from interface
// matrices D = (1/t) [[1, r], [r, 1]] and propagation matrices
// P = diag(exp(-i phi), exp(i phi)), phi = n (2 pi / lambda) d cos theta
// (scene/material.h:263-309, core/utils.h:173-193).  Because every factor is
// symmetric, M01/M00 is the reflection coefficient of the same stack lit from
// the far side.  That quantity is computed here with optical admittances,
// starting from the near medium and climbing layer by layer:
//   eta_s = n cos theta, eta_p = cos theta / n   (so r = (eta_a - eta_b) /
//   (eta_a + eta_b) reproduces Lambda's r_s and r_p),
//   Y <- eta (Y cos delta - i eta sin delta) / (eta cos delta - i Y sin delta),
//   r = (eta_far - Y) / (eta_far + Y).
// The two forms are algebraically identical (checked against an FP64
// restatement of Lambda's matrices on random absorbing multilayers).
//
// Precision: near grazing incidence a lossless film is a sharp Fabry-Perot
// resonance -- R swings from 0 to 1 while sin(delta) moves by about
// 2 cos(theta) / n -- so the phase (up to ~20 rad) and the recursion need far
// more than FP32's 24 bits to stay within 1e-5 of the FP64 value.  Both run
// in float-float arithmetic (ffloat.h); angles follow Lambda's Snell
// recurrence sin^2_{j+1} = (n_j / n_{j+1})^2 sin^2_j with principal complex
// square roots, so total internal reflection takes the same branch.

// Reflectance of thin-film material mt = {type, layers, offset of n, offset
// of d} at wavelength lambda, for a ray of direction d hitting a surface with
// normal n (facing the ray).  thick[] holds the layer thicknesses (nm) in
// material order; front selects the stack orientation (material.h:264-268).
LAM_OUTLINE float thinfilm_reflectance_at(const float2 *cspectra, const int *props, int4 mt, bool front, V3 n, V3 d,
                                          float lambda, const float *thick) {
    const int layers = mt.y, n_off = mt.z;
    const int j_off = front ? 0 : layers + 1;
    const int j_sign = front ? 1 : -1;
    auto index_of = [&](int j) {
        float2 v = dense_c(cspectra + size_t(props[n_off + j_off + j_sign * j]) * kBins, lambda);
        return cff(v.x, v.y);
    };
    const CFF one = cff(1.0f);
    // Snell recurrence, outer medium first.
    CFF nj = index_of(0);
    CFF cj = cff(dot(n, -d));
    CFF ys = nj * cj, yp = cj / nj;
    CFF n_next = index_of(1);
    CFF ratio = nj / n_next;
    CFF sin2 = ratio * (ratio * (one - cj * cj));
    CFF c_next = cffsqrt(one - sin2);
    // 2 pi / lambda rounded to float, as Lambda forms it (material.h:293)
    const FF k = ff(2 * kPi / lambda);
    for (int j = 1; j <= layers; j++) {
        nj = n_next;
        cj = c_next;
        n_next = index_of(j + 1);
        ratio = nj / n_next;
        sin2 = ratio * (ratio * sin2);
        c_next = cffsqrt(one - sin2);
        const FF dj = ff(thick[front ? j - 1 : layers - j]);
        const CFF delta = ((nj * k) * dj) * cj;
        // cos and sin of the complex phase, both divided by cosh(Im delta)
        FF sa, ca;
        ffsincos(delta.re, sa, ca);
        // a lossless layer has a real phase: tanh(0) = 0 exactly, so its
        // exp series is skipped
        const FF th = delta.im.hi == 0 && delta.im.lo == 0 ? ff(0.0f) : fftanh(delta.im);
        const CFF cd = CFF{ca, -(sa * th)};
        const CFF isd = CFF{-(ca * th), sa};  // i * sin(delta) / cosh
        const CFF es = nj * cj, ep = cj / nj;
        ys = es * (ys * cd - es * isd) / (es * cd - ys * isd);
        yp = ep * (yp * cd - ep * isd) / (ep * cd - yp * isd);
    }
    const CFF es_far = n_next * c_next, ep_far = c_next / n_next;
    const FF rs = cffabs2(es_far - ys) / cffabs2(es_far + ys);
    const FF rp = cffabs2(ep_far - yp) / cffabs2(ep_far + yp);
    return ffloat((rs + rp) * 0.5f);
}

LAM_FN float thinfilm_reflectance(const DeviceScene &S, int4 mt, bool front, V3 n, V3 d, float lambda,
                                  const float *thick) {
    return thinfilm_reflectance_at(S.cspectra, S.props, mt, front, n, d, lambda, thick);
}

constexpr int kMaxLayers = 8;

}  // namespace lam
