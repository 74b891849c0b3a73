// GPU-side BVH build for scenes whose geometry changes between frames (no
// reference counterpart: the reference builds a SAH BVH on the host,
// scene/bvh.h:61-246).  Linear BVHs (Karras, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees", HPG 2012): 30-bit Morton
// codes of the primitive centroids within their group's centroid box,
// radix-sorted with the group in the high word, a binary radix tree per
// group, boxes and subtree sizes refitted bottom-up, and nodes numbered in
// depth-first order so that the result has the node convention of the flat
// scene arrays (left child at i + 1, right child at `index`, a leaf of one
// primitive has count 1).  Object groups (the object range of each
// instance) get one tree each; the instances get the instance-level tree.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <map>
#include <vector>

#include "lbvh.h"
#include "runtime.h"
#include "dmath.h"

namespace lam {
namespace {

struct Box {
    float lo[3], hi[3];
};

// Spreads the 10 low bits of v to every third bit.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t morton3(float x, float y, float z) {
    auto q = [](float t) { return uint32_t(fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f)); };
    return (spread3(q(x)) << 2) | (spread3(q(y)) << 1) | spread3(q(z));
}

// Local-space box of object k: sphere centre +- r; quad and triangle the
// min / max of their vertices (corner, corner + h, corner + v[, + both]).
__global__ void k_object_boxes(const int *type, const float *geom, int n, Box *box) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float *g = geom + size_t(k) * 12;
    Box b;
    if (type[k] == LAMBDA_OBJECT_SPHERE) {
        const float r = fabsf(g[3]);
        for (int a = 0; a < 3; a++) {
            b.lo[a] = g[a] - r;
            b.hi[a] = g[a] + r;
        }
    } else {
        const bool quad = type[k] == LAMBDA_OBJECT_QUAD;
        for (int a = 0; a < 3; a++) {
            const float c = g[a], h = c + g[3 + a], v = c + g[6 + a];
            float lo = fminf(c, fminf(h, v)), hi = fmaxf(c, fmaxf(h, v));
            if (quad) {
                const float hv = c + g[3 + a] + g[6 + a];
                lo = fminf(lo, hv);
                hi = fmaxf(hi, hv);
            }
            b.lo[a] = lo;
            b.hi[a] = hi;
        }
    }
    box[k] = b;
}

// World box of instance i: its object-level root box through the
// object-to-world matrix (8 corners, homogeneous divide as in the traversal).
__global__ void k_instance_boxes(const float *mats, const Box *root_box, const int *group_of_instance, int n,
                                 Box *box) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *M = mats + size_t(i) * 32;
    const Box rb = root_box[group_of_instance[i]];
    Box b;
    for (int a = 0; a < 3; a++) {
        b.lo[a] = INFINITY;
        b.hi[a] = -INFINITY;
    }
    for (int c = 0; c < 8; c++) {
        const float p[3] = {(c & 1) ? rb.hi[0] : rb.lo[0], (c & 2) ? rb.hi[1] : rb.lo[1], (c & 4) ? rb.hi[2] : rb.lo[2]};
        float w = M[12] * p[0] + M[13] * p[1] + M[14] * p[2] + M[15];
        if (fabsf(w) < 1e-4f) w = 1;
        for (int a = 0; a < 3; a++) {
            const float x = (M[4 * a] * p[0] + M[4 * a + 1] * p[1] + M[4 * a + 2] * p[2] + M[4 * a + 3]) / w;
            b.lo[a] = fminf(b.lo[a], x);
            b.hi[a] = fmaxf(b.hi[a], x);
        }
    }
    // widen by a few ulps: the traversal transforms rays, not boxes
    for (int a = 0; a < 3; a++) {
        const float e = 1e-6f * fmaxf(fabsf(b.lo[a]), fabsf(b.hi[a])) + 1e-30f;
        b.lo[a] -= e;
        b.hi[a] += e;
    }
    box[i] = b;
}

// Centroid box of each group (one block per group over its range).
__global__ void __launch_bounds__(256) k_group_centroid_box(const Box *box, const int *gfirst, const int *gcount,
                                                            Box *cbox) {
    __shared__ float red[6][256];
    const int g = blockIdx.x;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = gfirst[g] + threadIdx.x; k < gfirst[g] + gcount[g]; k += 256)
        for (int a = 0; a < 3; a++) {
            const float c = 0.5f * (box[k].lo[a] + box[k].hi[a]);
            lo[a] = fminf(lo[a], c);
            hi[a] = fmaxf(hi[a], c);
        }
    for (int a = 0; a < 3; a++) {
        red[a][threadIdx.x] = lo[a];
        red[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off)
            for (int a = 0; a < 3; a++) {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + off]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + off]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; a++) {
            cbox[g].lo[a] = red[a][0];
            cbox[g].hi[a] = red[3 + a][0];
        }
}

// key = group << 32 | Morton code of the centroid in the group's centroid box.
__global__ void k_morton(const Box *box, const int *group, const Box *cbox, int n, unsigned long long *key,
                         int *val) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Box c = cbox[group[k]];
    float t[3];
    for (int a = 0; a < 3; a++) {
        const float ext = c.hi[a] - c.lo[a];
        const float x = 0.5f * (box[k].lo[a] + box[k].hi[a]);
        t[a] = ext > 0 ? (x - c.lo[a]) / ext : 0.5f;
    }
    key[k] = (unsigned long long)uint32_t(group[k]) << 32 | morton3(t[0], t[1], t[2]);
    val[k] = k;
}

// Common-prefix length of sorted keys i and j of one group (positions local
// to the group, -1 outside it); equal codes compare their positions.
__device__ __forceinline__ int delta(const unsigned long long *key, int first, int m, int i, int j) {
    if (j < 0 || j >= m) return -1;
    const uint32_t a = uint32_t(key[first + i]), b = uint32_t(key[first + j]);
    return a != b ? __clz(a ^ b) : 32 + __clz(uint32_t(i ^ j));
}

// Binary radix tree of every group with at least two primitives: internal
// node i (local) of group g is global node inner_base[g] + i; its children
// are leaves (>= 0 flag bit 30 set) or internal nodes.  child[2 * node + s],
// parent of internal nodes and leaves (leaves: parent_leaf[position]).
constexpr int kLeafBit = 1 << 30;

__global__ void k_karras(const unsigned long long *key, const int *gfirst, const int *gcount, const int *gof_pos,
                         const int *inner_base, int n, int *child, int *parent_inner, int *parent_leaf) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int g = gof_pos[p];
    const int first = gfirst[g], m = gcount[g];
    const int i = p - first;
    if (i >= m - 1) return;  // m - 1 internal nodes
    const int d = delta(key, first, m, i, i + 1) > delta(key, first, m, i, i - 1) ? 1 : -1;
    const int dmin = delta(key, first, m, i, i - d);
    int lmax = 2;
    while (delta(key, first, m, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(key, first, m, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(key, first, m, i, j);
    int s = 0;
    for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
        if (delta(key, first, m, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    const int split = i + s * d + min(d, 0);
    const int base = inner_base[g];
    const int node = base + i;
    const int left = min(i, j) == split ? (kLeafBit | (first + split)) : base + split;
    const int right = max(i, j) == split + 1 ? (kLeafBit | (first + split + 1)) : base + split + 1;
    child[2 * node] = left;
    child[2 * node + 1] = right;
    if (left & kLeafBit) parent_leaf[left & ~kLeafBit] = node;
    else parent_inner[left] = node;
    if (right & kLeafBit) parent_leaf[right & ~kLeafBit] = node;
    else parent_inner[right] = node;
}

// Bottom-up boxes and subtree sizes (nodes): each leaf climbs; the second
// child to arrive at an internal node finishes it and climbs on.
__global__ void k_refit(const Box *leaf_box, const int *child, const int *parent_inner, const int *parent_leaf,
                        const int *gof_pos, const int *gcount, int n, int *arrive, Box *inner_box, int *inner_size) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || gcount[gof_pos[p]] < 2) return;
    int node = parent_leaf[p];
    while (node >= 0) {
        __threadfence();
        if (atomicAdd(arrive + node, 1) == 0) return;
        __threadfence();
        Box b;
        int size = 1;
        for (int a = 0; a < 3; a++) {
            b.lo[a] = INFINITY;
            b.hi[a] = -INFINITY;
        }
        for (int s = 0; s < 2; s++) {
            const int c = child[2 * node + s];
            Box cb;
            if (c & kLeafBit) {
                cb = leaf_box[c & ~kLeafBit];
                size += 1;
            } else {
                cb = inner_box[c];
                size += inner_size[c];
            }
            for (int a = 0; a < 3; a++) {
                b.lo[a] = fminf(b.lo[a], cb.lo[a]);
                b.hi[a] = fmaxf(b.hi[a], cb.hi[a]);
            }
        }
        inner_box[node] = b;
        inner_size[node] = size;
        node = parent_inner[node];
    }
}

// Depth-first numbering: each node sums, up its ancestors, 1 for being a
// child plus the left sibling's subtree size when it is a right child; then
// writes its flat record at out_base[g] + number.
__global__ void k_flatten(const Box *leaf_box, const int *child, const int *parent_inner, const int *parent_leaf,
                          const Box *inner_box, const int *inner_size, const int *gof_pos, const int *gof_inner,
                          const int *out_base, int n, int n_inner, float *node_bounds, int *node_ints, int prim_base) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    // elements: leaves p in [0, n), then internal nodes n + node
    if (e >= n + n_inner) return;
    const bool leaf = e < n;
    const int g = leaf ? gof_pos[e] : gof_inner[e - n];
    const int self = leaf ? (kLeafBit | e) : e - n;
    int number = 0;
    int x = self;
    while (true) {
        const int par = (x & kLeafBit) ? parent_leaf[x & ~kLeafBit] : parent_inner[x];
        if (par < 0) break;
        number += 1;
        if (child[2 * par + 1] == x) {
            const int l = child[2 * par];
            number += (l & kLeafBit) ? 1 : inner_size[l];
        }
        x = par;
    }
    const int o = out_base[g] + number;
    Box b;
    if (leaf) {
        b = leaf_box[e];
        node_ints[2 * o] = prim_base + e;
        node_ints[2 * o + 1] = 1;
    } else {
        b = inner_box[self];
        // the right child's number: this node's number + 1 + the left size
        const int l = child[2 * self];
        node_ints[2 * o] = out_base[g] + number + 1 + ((l & kLeafBit) ? 1 : inner_size[l]);
        node_ints[2 * o + 1] = 0;
    }
    for (int a = 0; a < 3; a++) {
        node_bounds[6 * o + a] = b.lo[a];
        node_bounds[6 * o + 3 + a] = b.hi[a];
    }
}

__global__ void k_gather_box(const Box *src, const int *idx, int n, Box *dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[idx[k]];
}

int blocks(long long n) { return int(std::max(1LL, (n + 255) / 256)); }

// Builds one linear BVH per group over the boxes of `n` primitives (groups
// tile [0, n) in order); order[p] = source primitive of sorted position p.
// Node records of group g go to [out_base[g], out_base[g] + 2 count - 1).
void build_forest(const Box *d_box, int n, const std::vector<int> &gfirst, const std::vector<int> &gcount,
                  const std::vector<int> &out_base, int prim_base, std::vector<int> &order, float *h_bounds,
                  int *h_ints, int total_nodes, Box *root_box_out) {
    const int G = int(gfirst.size());
    // internal nodes of group g: inner_base[g] + i, i < count - 1 (contiguous)
    std::vector<int> gof(static_cast<size_t>(n)), inner_base(static_cast<size_t>(G)), gof_inner;
    for (int g = 0; g < G; g++) {
        inner_base[size_t(g)] = gfirst[size_t(g)] - g;
        for (int k = 0; k < gcount[size_t(g)]; k++) gof[size_t(gfirst[size_t(g)] + k)] = g;
        for (int k = 0; k + 1 < gcount[size_t(g)]; k++) gof_inner.push_back(g);
    }
    const int n_inner = int(gof_inner.size());
    DevBuf<int> d_gof_inner;
    if (n_inner) d_gof_inner.upload(gof_inner);
    DevBuf<int> d_gfirst, d_gcount, d_gof, d_inner_base, d_out_base;
    d_gfirst.upload(gfirst);
    d_gcount.upload(gcount);
    d_gof.upload(gof);
    d_inner_base.upload(inner_base);
    d_out_base.upload(out_base);
    DevBuf<Box> cbox, sorted_box, inner_box;
    cbox.alloc(size_t(G));
    k_group_centroid_box<<<G, 256>>>(d_box, d_gfirst.p, d_gcount.p, cbox.p);
    DevBuf<unsigned long long> key, key_sorted;
    DevBuf<int> val, val_sorted;
    key.alloc(size_t(n));
    key_sorted.alloc(size_t(n));
    val.alloc(size_t(n));
    val_sorted.alloc(size_t(n));
    k_morton<<<blocks(n), 256>>>(d_box, d_gof.p, cbox.p, n, key.p, val.p);
    HIP_OK(hipGetLastError());
    size_t temp = 0;
    HIP_OK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, key.p, key_sorted.p, val.p, val_sorted.p, n));
    DevBuf<char> tmp;
    tmp.alloc(std::max<size_t>(temp, 1));
    HIP_OK(hipcub::DeviceRadixSort::SortPairs(tmp.p, temp, key.p, key_sorted.p, val.p, val_sorted.p, n));
    sorted_box.alloc(size_t(n));
    k_gather_box<<<blocks(n), 256>>>(d_box, val_sorted.p, n, sorted_box.p);
    DevBuf<int> child, parent_inner, parent_leaf, arrive, inner_size;
    const size_t ni = size_t(std::max(n, 1));
    child.alloc(2 * ni);
    parent_inner.alloc(ni);
    parent_leaf.alloc(ni);
    arrive.alloc(ni);
    inner_size.alloc(ni);
    inner_box.alloc(ni);
    HIP_OK(hipMemset(parent_inner.p, 0xff, ni * sizeof(int)));
    HIP_OK(hipMemset(parent_leaf.p, 0xff, ni * sizeof(int)));
    HIP_OK(hipMemset(arrive.p, 0, ni * sizeof(int)));
    k_karras<<<blocks(n), 256>>>(key_sorted.p, d_gfirst.p, d_gcount.p, d_gof.p, d_inner_base.p, n, child.p,
                                 parent_inner.p, parent_leaf.p);
    k_refit<<<blocks(n), 256>>>(sorted_box.p, child.p, parent_inner.p, parent_leaf.p, d_gof.p, d_gcount.p, n, arrive.p,
                                inner_box.p, inner_size.p);
    DevBuf<float> d_bounds;
    DevBuf<int> d_ints;
    d_bounds.alloc(size_t(total_nodes) * 6);
    d_ints.alloc(size_t(total_nodes) * 2);
    k_flatten<<<blocks((long long)n + n_inner), 256>>>(sorted_box.p, child.p, parent_inner.p, parent_leaf.p,
                                                       inner_box.p, inner_size.p, d_gof.p, d_gof_inner.p, d_out_base.p,
                                                       n, n_inner, d_bounds.p, d_ints.p, prim_base);
    HIP_OK(hipGetLastError());
    order.resize(size_t(n));
    HIP_OK(hipMemcpy(order.data(), val_sorted.p, size_t(n) * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<float> b(size_t(total_nodes) * 6);
    std::vector<int> in(size_t(total_nodes) * 2);
    HIP_OK(hipMemcpy(b.data(), d_bounds.p, b.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(in.data(), d_ints.p, in.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::copy(b.begin(), b.end(), h_bounds);
    std::copy(in.begin(), in.end(), h_ints);
    if (root_box_out)
        for (int g = 0; g < G; g++) {
            const float *r = h_bounds + size_t(out_base[size_t(g)]) * 6;
            for (int a = 0; a < 3; a++) {
                root_box_out[g].lo[a] = r[a];
                root_box_out[g].hi[a] = r[3 + a];
            }
        }
}

}  // namespace

void build_lbvh(const lambda_scene_view &v, LbvhScene &out) {
    const int O = v.num_objects, I = v.num_instances;
    if (O <= 0 || I <= 0) throw Failure{LAMBDA_EINVAL, "invalid scene: empty scene"};
    // object groups: the distinct object ranges of the instances, which must
    // tile the object array
    std::map<int, int> range;  // first -> count
    for (int i = 0; i < I; i++) {
        const int f = v.instance_ints[4 * i], c = v.instance_ints[4 * i + 1];
        if (f < 0 || c < 1 || (long long)f + c > O) throw Failure{LAMBDA_EINVAL, "invalid scene: instance object range"};
        auto it = range.find(f);
        if (it != range.end() && it->second != c)
            throw Failure{LAMBDA_EINVAL, "GPU BVH build: instances share objects with different ranges"};
        range[f] = c;
    }
    std::vector<int> gfirst, gcount;
    int next = 0;
    for (auto &rc : range) {
        if (rc.first != next) throw Failure{LAMBDA_EINVAL, "GPU BVH build: instance object ranges must tile the objects"};
        gfirst.push_back(rc.first);
        gcount.push_back(rc.second);
        next = rc.first + rc.second;
    }
    if (next != O) throw Failure{LAMBDA_EINVAL, "GPU BVH build: instance object ranges must tile the objects"};
    const int G = int(gfirst.size());
    std::map<int, int> group_of_first;
    for (int g = 0; g < G; g++) group_of_first[gfirst[size_t(g)]] = g;

    // node layout: the instance level first (2 I - 1 nodes), then each group
    const int tlas_nodes = 2 * I - 1;
    std::vector<int> blas_base(static_cast<size_t>(G));
    int total = tlas_nodes;
    for (int g = 0; g < G; g++) {
        blas_base[size_t(g)] = total;
        total += 2 * gcount[size_t(g)] - 1;
    }
    out.node_bounds.assign(size_t(total) * 6, 0.0f);
    out.node_ints.assign(size_t(total) * 2, 0);

    // object level
    DevBuf<int> d_type;
    DevBuf<float> d_geom;
    d_type.upload(v.object_type, size_t(O));
    d_geom.upload(v.object_geometry, size_t(O) * 12);
    DevBuf<Box> obox;
    obox.alloc(size_t(O));
    k_object_boxes<<<blocks(O), 256>>>(d_type.p, d_geom.p, O, obox.p);
    HIP_OK(hipGetLastError());
    std::vector<int> obj_order;
    std::vector<Box> root(static_cast<size_t>(G));
    std::vector<int> blas_local(blas_base.size());
    for (int g = 0; g < G; g++) blas_local[size_t(g)] = blas_base[size_t(g)] - tlas_nodes;
    build_forest(obox.p, O, gfirst, gcount, blas_local, 0, obj_order, out.node_bounds.data() + size_t(tlas_nodes) * 6,
                 out.node_ints.data() + size_t(tlas_nodes) * 2, total - tlas_nodes, root.data());
    // object-level indices point into the whole node array
    for (int k = tlas_nodes; k < total; k++)
        if (out.node_ints[size_t(k) * 2 + 1] == 0) out.node_ints[size_t(k) * 2] += tlas_nodes;

    // instance level over the instances' world boxes
    std::vector<int> inst_group(static_cast<size_t>(I));
    for (int i = 0; i < I; i++) inst_group[size_t(i)] = group_of_first[v.instance_ints[4 * i]];
    DevBuf<float> d_mats;
    d_mats.upload(v.instance_matrices, size_t(I) * 32);
    DevBuf<Box> d_root, ibox;
    d_root.upload(root);
    DevBuf<int> d_ig;
    d_ig.upload(inst_group);
    ibox.alloc(size_t(I));
    k_instance_boxes<<<blocks(I), 256>>>(d_mats.p, d_root.p, d_ig.p, I, ibox.p);
    HIP_OK(hipGetLastError());
    std::vector<int> inst_order;
    build_forest(ibox.p, I, std::vector<int>{0}, std::vector<int>{I}, std::vector<int>{0}, 0, inst_order,
                 out.node_bounds.data(), out.node_ints.data(), tlas_nodes, nullptr);
    out.object_order = obj_order;
    out.instance_order = inst_order;
    out.blas_root.resize(size_t(I));
    for (int i = 0; i < I; i++) out.blas_root[size_t(i)] = blas_base[size_t(inst_group[size_t(i)])];
}

}  // namespace lam
