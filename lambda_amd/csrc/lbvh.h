// GPU-side linear BVH build (lbvh.hip): node arrays in the convention of
// the flat scene arrays, and the reordering of objects and instances.
#pragma once

#include <vector>

#include "lambda_hip.h"

namespace lam {

struct LbvhScene {
    std::vector<float> node_bounds;   // [nodes][6]
    std::vector<int> node_ints;       // [nodes][2] = {index, count}
    std::vector<int> object_order;    // new object k = source object object_order[k]
    std::vector<int> instance_order;  // new instance i = source instance instance_order[i]
    std::vector<int> blas_root;       // object-level root of each source instance
};

// Builds the BVHs of v's objects and instances on the current device (v's
// node arrays are not read).
void build_lbvh(const lambda_scene_view &v, LbvhScene &out);

}  // namespace lam
