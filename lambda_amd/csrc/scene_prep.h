// Host-side preparation of the device scene: packing of objects and
// instances into their device records, the 4-wide BVH built over the
// reference's nodes, and the reference's depth-first ranks used to break
// exact distance ties.  Plain C++ on lambda_scene_view; shared by the HIP
// library's upload and by host-compiled checks of the traversal.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "lambda_hip.h"

namespace lam {

// obj[4 n .. 4 n + 3]: {corner | centre, radius}, {horizontal, bits(type)},
// {vertical, 0}, {normal, 0}.
inline std::vector<float4> pack_objects(const lambda_scene_view &v) {
    const size_t O = size_t(v.num_objects);
    std::vector<float4> out(4 * O);
    for (size_t i = 0; i < O; i++) {
        const float *g0 = v.object_geometry + 12 * i;
        int type = v.object_type[i];
        float ft;
        std::memcpy(&ft, &type, 4);
        float4 *r = &out[4 * i];
        if (type == LAMBDA_OBJECT_SPHERE) {
            r[0] = make_float4(g0[0], g0[1], g0[2], g0[3]);
            r[1] = make_float4(0, 0, 0, ft);
            r[2] = make_float4(0, 0, 0, 0);
            r[3] = make_float4(0, 0, 0, 0);
        } else {
            r[0] = make_float4(g0[0], g0[1], g0[2], 0);
            r[1] = make_float4(g0[3], g0[4], g0[5], ft);
            r[2] = make_float4(g0[6], g0[7], g0[8], 0);
            r[3] = make_float4(g0[9], g0[10], g0[11], 0);
        }
    }
    return out;
}

// obj_s[4 n .. 4 n + 3]: the 15 shading floats of a triangle, padded.
inline std::vector<float4> pack_shading(const lambda_scene_view &v) {
    const size_t O = size_t(v.num_objects);
    std::vector<float4> os(4 * O);
    for (size_t i = 0; i < O; i++) {
        const float *sh = v.object_shading + 15 * i;
        os[4 * i + 0] = make_float4(sh[0], sh[1], sh[2], sh[3]);
        os[4 * i + 1] = make_float4(sh[4], sh[5], sh[6], sh[7]);
        os[4 * i + 2] = make_float4(sh[8], sh[9], sh[10], sh[11]);
        os[4 * i + 3] = make_float4(sh[12], sh[13], sh[14], 0);
    }
    return os;
}

// inst[8 i ..]: object-to-world rows, then world-to-object rows;
// inst_i[i] = {first object, count, BLAS root, material}.
inline void pack_instances(const lambda_scene_view &v, std::vector<float4> &im, std::vector<int4> &ii) {
    const size_t I = size_t(v.num_instances);
    im.resize(8 * I);
    ii.resize(I);
    for (size_t i = 0; i < I; i++) {
        const float *m = v.instance_matrices + 32 * i;
        for (int r = 0; r < 8; r++) im[8 * i + r] = make_float4(m[4 * r], m[4 * r + 1], m[4 * r + 2], m[4 * r + 3]);
        const int *q = v.instance_ints + 4 * i;
        ii[i] = make_int4(q[0], q[1], q[2], q[3]);
    }
}

// Replays the reference's depth-first order (left pushed before right, right
// popped first; scene/scene.h:112-119) without culling and numbers the leaves
// in the order it meets them: tlas_rank[instance] over the instance-level
// tree, obj_rank[object] within its object-level tree.
inline void reference_ranks(const lambda_scene_view &v, std::vector<int> &tlas_rank, std::vector<int> &obj_rank) {
    const size_t I = size_t(v.num_instances), O = size_t(v.num_objects);
    tlas_rank.assign(I, 0x7fffffff);
    obj_rank.assign(O, 0x7fffffff);
    auto replay = [&](int root, auto &&on_leaf) {
        std::vector<int> st{root};
        while (!st.empty()) {
            int n = st.back();
            st.pop_back();
            int idx = v.node_ints[2 * n], cnt = v.node_ints[2 * n + 1];
            if (cnt > 0) on_leaf(idx, cnt);
            else {
                st.push_back(n + 1);
                st.push_back(idx);
            }
        }
    };
    int r = 0;
    replay(0, [&](int idx, int) { tlas_rank[size_t(idx)] = r++; });
    std::vector<char> seen(size_t(v.num_nodes), 0);
    for (size_t i = 0; i < I; i++) {
        int root = v.instance_ints[4 * i + 2];
        if (seen[size_t(root)]) continue;
        seen[size_t(root)] = 1;
        int k = 0;
        replay(root, [&](int idx, int cnt) {
            for (int j = 0; j < cnt; j++) obj_rank[size_t(idx + j)] = k++;
        });
    }
}

// 4-wide BVH over the reference's nodes.  A wide node stands for one
// reference node X and holds up to four of X's descendants: starting from
// X's two children, the inner child of largest surface area is replaced by
// its two children until there are four or only leaves remain.  A leaf
// reference node (the root of a one-node tree) is a wide node with one leaf
// slot.  need[] is the deepest traversal stack a subtree can build: a wide
// node pushes at most (descendable slots - 1) entries before entering one
// of them (descendable: inner slots, and instance leaves at the instance
// level, whose object-level trees count as their subtrees).
struct WideBuilder {
    const lambda_scene_view &v;
    std::vector<float4> nodes;   // 8 per wide node
    std::vector<int> need;
    std::vector<int> memo;       // reference node -> wide node of an object-level root, or -1
    std::vector<int> inst_root;  // [instances] wide root
    std::vector<int> inst_need;

    explicit WideBuilder(const lambda_scene_view &view)
        : v(view), memo(size_t(view.num_nodes), -1) {}

    float area(int n) const {
        const float *b = v.node_bounds + 6 * size_t(n);
        float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
        return 2 * (dx * dy + dy * dz + dz * dx);
    }
    bool leaf(int n) const { return v.node_ints[2 * size_t(n) + 1] > 0; }

    int build(int x, bool instance_level) {
        std::vector<int> slots;
        if (leaf(x)) {
            slots.push_back(x);
        } else {
            slots.push_back(x + 1);
            slots.push_back(v.node_ints[2 * size_t(x)]);
            while (slots.size() < 4) {
                int pick = -1;
                for (int j = 0; j < int(slots.size()); j++)
                    if (!leaf(slots[size_t(j)]) && (pick < 0 || area(slots[size_t(j)]) > area(slots[size_t(pick)])))
                        pick = j;
                if (pick < 0) break;
                const int n = slots[size_t(pick)];
                slots[size_t(pick)] = n + 1;
                slots.push_back(v.node_ints[2 * size_t(n)]);
            }
        }
        const int w = int(nodes.size() / 8);
        nodes.resize(nodes.size() + 8);
        need.push_back(0);
        float lo[3][4], hi[3][4];
        int child[4], count[4];
        int descend = 0, deepest = 0;
        for (int j = 0; j < 4; j++) {
            if (j >= int(slots.size())) {
                for (int a = 0; a < 3; a++) lo[a][j] = hi[a][j] = 0;
                child[j] = 0;
                count[j] = -1;
                continue;
            }
            const int n = slots[size_t(j)];
            const float *b = v.node_bounds + 6 * size_t(n);
            for (int a = 0; a < 3; a++) {
                lo[a][j] = b[a];
                hi[a][j] = b[3 + a];
            }
            if (leaf(n)) {
                child[j] = v.node_ints[2 * size_t(n)];
                count[j] = v.node_ints[2 * size_t(n) + 1];
                if (instance_level) {
                    descend++;
                    deepest = std::max(deepest, inst_need[size_t(child[j])]);
                }
            } else {
                child[j] = build(n, instance_level);
                count[j] = 0;
                descend++;
                deepest = std::max(deepest, need[size_t(child[j])]);
            }
        }
        need[size_t(w)] = descend > 0 ? descend - 1 + deepest : 0;
        float4 *o = &nodes[8 * size_t(w)];
        for (int a = 0; a < 3; a++) {
            o[2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
            o[2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
        }
        float fc[4], fn[4];
        std::memcpy(fc, child, sizeof fc);
        std::memcpy(fn, count, sizeof fn);
        o[6] = make_float4(fc[0], fc[1], fc[2], fc[3]);
        o[7] = make_float4(fn[0], fn[1], fn[2], fn[3]);
        return w;
    }

    // Returns the stack depth bound of the whole scene.
    int run() {
        // the instance-level tree is wide node 0; object-level trees follow
        const size_t I = size_t(v.num_instances);
        inst_root.assign(I, 0);
        inst_need.assign(I, 0);
        nodes.resize(8);
        need.push_back(0);
        for (size_t i = 0; i < I; i++) {
            const int r = v.instance_ints[4 * i + 2];
            if (memo[size_t(r)] < 0) memo[size_t(r)] = build(r, false);
            inst_root[i] = memo[size_t(r)];
            inst_need[i] = need[size_t(inst_root[i])];
        }
        // build the instance level into a scratch builder so that it lands at 0
        std::vector<float4> blas = std::move(nodes);
        std::vector<int> blas_need = std::move(need);
        nodes.clear();
        need.clear();
        const int root = build(0, true);
        (void)root;
        const int top = int(nodes.size() / 8);
        const int total = need[0];
        // append the object-level nodes after the instance level, shifting
        // their indices (wide node 0 of `blas` was a placeholder)
        const int shift = top - 1;
        for (size_t w = 1; w < blas.size() / 8; w++) {
            float4 *o = &blas[8 * w];
            int c[4], n[4];
            std::memcpy(c, &o[6], sizeof c);
            std::memcpy(n, &o[7], sizeof n);
            for (int j = 0; j < 4; j++)
                if (n[j] == 0) c[j] += shift;
            std::memcpy(&o[6], c, sizeof c);
            nodes.insert(nodes.end(), o, o + 8);
        }
        for (auto &r : inst_root) r += shift;
        (void)blas_need;
        return total;
    }
};

}  // namespace lam
