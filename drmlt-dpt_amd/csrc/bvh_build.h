// Host BVH builder interface (binned SAH, depth-capped, collapsed to 4-wide nodes).
#pragma once
#include <vector>

namespace drmlt {

struct BvhOut {
    std::vector<float> nodes;  // 32 floats per 4-wide node (see bvh_build.cpp)
    std::vector<int> order;    // leaf-order position -> scene-order triangle index
    int depth = 0;             // binary tree depth
    int depth4 = 0;            // 4-wide tree depth
    int stackNeed = 0;         // largest traversal stack a ray can need
};

int BuildBvh(int nt, const float *triP, int maxDepth, BvhOut &out);

}  // namespace drmlt
