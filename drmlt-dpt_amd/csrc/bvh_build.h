// Host BVH builder interface (binned SAH, depth-capped).
#pragma once
#include <vector>

namespace drmlt {

struct BvhOut {
    std::vector<float> nodes;  // 16 floats per internal node (see bvh_build.cpp)
    std::vector<int> order;    // leaf-order position -> scene-order triangle index
    int depth = 0;
};

int BuildBvh(int nt, const float *triP, int maxDepth, BvhOut &out);

}  // namespace drmlt
