// The reference's serialized forms of a subpath and of its scene, the
// arguments of its derivative functions (PathFuncDerv, src/path.h:145-149),
// read into this library's path records. Host code.
//
// Subpath (Serialize, src/path.cpp:2303-2391): `primary` holds the time,
// then the primary sample in GetPathPos order; `vertParams` holds, in order,
// the lens vertex position (3); for a light subpath of two or more vertices
// the light pick probability (1), the emitting light (a light record) and per
// light surface vertex a shape record, bsdfDiscrete, useAbsoluteParam, a
// BSDF record and (all but the last) the Russian-roulette weight; then per
// camera surface vertex a shape record and, except at the last, bsdfDiscrete,
// useAbsoluteParam, a BSDF record and the Russian-roulette weight. The last
// camera vertex is followed by the light it hits and the pick probability
// (no light subpath), by the direct light, its BSDF record and the pick
// probability (one light vertex), or by its BSDF record.
//
// Records are fixed-size, padded to the largest kind: a shape (triangle
// mesh, src/trianglemesh.cpp:3-11, 143-185) is type, isMoving, then per time
// p0, p1 - p0, p2 - p0, n0, n1, n2, then the no-uv flag, three uvs and the
// inverse mesh area (46 floats); a BSDF (src/bsdf.cpp:7-11) is type and its
// parameters (10); a light (src/light.cpp:7-10) is type and its parameters,
// for an area light (src/arealight.cpp:6-22) a shape record and the emission
// (56). The scene (src/scene.cpp:122-131) is useLightCoordinateSampling, the
// camera (src/camera.cpp:4-36: sampleToCam column-major, camToWorld as
// isMoving + two translations + two quaternions, pixel count, dist) and the
// bounding sphere (center, radius): 38 floats.
#pragma once
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/drmlt_amd.h"

namespace drmlt {

constexpr int kSerShape = 46;
constexpr int kSerBsdf = 10;
constexpr int kSerLight = 56;
constexpr int kSerCamera = 33;
constexpr int kSerScene = 38;

// path.cpp:2286-2289 (time included)
inline int ser_primary_size(int c, int l) { return 2 * (c + l - 1 > 2 ? c + l - 1 : 2) + 1; }
// path.cpp:2291-2301
inline int64_t ser_vert_size(int maxC, int maxL) {
    const int64_t m = maxC + maxL;
    return m * kSerShape + m * kSerBsdf + kSerLight + 2 * m + m + 3 + 1 + 1;
}

// Scene-order triangles keyed by the exact bits of (p0, p1 - p0, p2 - p0)
// as a serialized shape carries them (time 0).
class TriKeyMap {
  public:
    // p0e1e2: nt x 9 floats
    void build(int nt, const float *p0e1e2);
    // the triangle of a shape record, -1 if none
    int find(const float *shape) const;
    bool empty() const { return map_.empty(); }

  private:
    struct Key {
        uint32_t w[9];
        bool operator==(const Key &o) const;
    };
    struct KeyHash {
        size_t operator()(const Key &k) const;
    };
    std::unordered_map<Key, int, KeyHash> map_;
};

// What a decoded state is checked against: the type of each triangle's
// material and each triangle's area light (-1: none).
struct SceneFacts {
    const int *triMatType;  // nt
    const int *triLight;    // nt
};

// One serialized state of (camera, light) vertex counts (c, l) into a path
// record: c, l, light and its triangle, screen position (lens), u (primary
// after the time), triangles, bsdfDiscrete and Russian-roulette weights;
// fields the layout does not carry (bsdfDiscrete of the last camera vertex,
// roulette weights of the last vertex of each subpath, slots past the path)
// take the record defaults (0, 1, triangle -1). Returns 0, or -1 with err.
int decode_subpath(const TriKeyMap &tm, const SceneFacts &sf, int c, int l, const float *lens, const float *primary, const float *vert,
                   int64_t vertLen, drmlt_path_rec &R, std::string &err);

// A serialized BSDF record (type + parameters: Lambertian Kd; Phong Kd, Ks,
// exponent, KsWeight; RoughDielectric Ks, Kt, eta, invEta, alpha;
// lambertian.cpp:10-13, phong.cpp:13-19, roughdielectric.cpp:13-20) as this
// library's 16-float material record. Returns false for an unknown type.
bool material_from_serialized(const float *bsdf, bool twoSided, float *mat);

}  // namespace drmlt
