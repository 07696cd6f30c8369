// Device BVH build (bvh_device.hip): the binned-SAH tree of bvh_build.cpp
// built on the GPU, written straight into the traversal's 4-wide node layout.
#pragma once
#include <hip/hip_runtime.h>

namespace drmlt {

struct DeviceBvh {
    float4 *nodes = nullptr;  // 8 float4 per 4-wide node (bvh_build.cpp layout); hipFree by the caller
    float4 *tri = nullptr;    // leaf-order triangles: p0 (.w = scene index bits), e1, e2; hipFree by the caller
    int numNodes = 0;
    int depth = 0;      // binary tree depth
    int depth4 = 0;     // 4-wide tree depth
    int stackNeed = 0;  // largest traversal stack a ray can need
    double ms = 0;      // build time (device work and the level loop's host waits)
};

// triP: nt x 9 floats on the device (the scene's triangle corners); s0, s1,
// s2: p0, p1 - p0, p2 - p0 per triangle on the device. maxDepth caps the
// binary tree as the host builder does. Returns hipSuccess, the first error,
// or hipErrorInvalidValue for a scene that is a single leaf (at most four
// triangles, or every centroid equal), which the host builder handles.
hipError_t BuildBvhDevice(int nt, const float *triP, const float4 *s0, const float4 *s1, const float4 *s2, int maxDepth, hipStream_t s,
                          DeviceBvh &out);

}  // namespace drmlt
