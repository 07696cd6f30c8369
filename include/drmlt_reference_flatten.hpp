// drmlt_reference_flatten.hpp -- the reference's in-memory Scene as the flat
// arrays of drmlt_scene_desc (include/drmlt_amd.h). Reference-side code: it
// includes the reference's own headers and compiles inside its source tree
// (src/), next to an integrator that calls the library.
//
// What it reads (the reference's types, src/scene.h:21-44):
//  * every Shape of scene->objects, a TriangleMesh (src/trianglemesh.h:19-68):
//    its triangles' time-0 positions and shading normals (TriMeshData
//    position0 / normal0 / indices; geometric normals where a mesh has
//    none), its BSDF and its area light;
//  * each BSDF through its own Serialize(st = (0, 0)) (the library supports
//    constant parameters) converted by drmlt_material_from_serialized, and
//    twoSided from Lambertian / Phong;
//  * each AreaLight of scene->lights (src/arealight.h): the mesh it sits on,
//    its emission and samplingWeight; the triangles of one light are one
//    mesh, contiguous in the flattened order;
//  * the Camera (src/camera.h:10-22): camToWorld at time 0 (the columns of
//    LookAt, src/transform.cpp:42-77: left, up, view direction, position),
//    the horizontal fov recovered from dist = W / (2 tan(fov / 2))
//    (src/camera.cpp:27), film size and clip planes.
// Only scenes the library renders flatten: triangle meshes, the three
// BSDFs, area lights and a pinhole camera; anything else throws.
#pragma once

#include <cmath>
#include <map>
#include <stdexcept>
#include <vector>

#include "arealight.h"
#include "bsdf.h"
#include "camera.h"
#include "image.h"
#include "lambertian.h"
#include "phong.h"
#include "scene.h"
#include "trianglemesh.h"
#include "transform.h"

#include "drmlt_amd.h"

namespace drmlt_ref {

struct FlatScene {
    std::vector<float> triP, triN, materials, lightEmit;
    std::vector<int32_t> triMat, triLight, lightTris;
    float camera[14];
    // the arrays as the library's scene description (valid while *this lives)
    drmlt_scene_desc desc() const {
        drmlt_scene_desc d;
        d.num_triangles = int32_t(triMat.size());
        d.tri_p = triP.data();
        d.tri_n = triN.data();
        d.tri_mat = triMat.data();
        d.tri_light = triLight.data();
        d.num_materials = int32_t(materials.size() / 16);
        d.materials = materials.data();
        d.num_lights = int32_t(lightTris.size() / 2);
        d.light_tris = lightTris.data();
        d.light_emit = lightEmit.data();
        d.camera = camera;
        return d;
    }
};

inline FlatScene Flatten(const Scene &scene) {
    FlatScene f;
    std::map<const BSDF *, int> matIndex;
    std::map<const AreaLight *, int> lightIndex;
    for (const auto &L : scene.lights) {
        const AreaLight *al = dynamic_cast<const AreaLight *>(L.get());
        if (!al) throw std::runtime_error("Flatten: only area lights are supported");
        const int k = int(lightIndex.size());
        lightIndex[al] = k;
    }
    f.lightTris.assign(2 * lightIndex.size(), -1);
    f.lightEmit.assign(4 * lightIndex.size(), 0.0f);
    for (const auto &obj : scene.objects) {
        const TriangleMesh *mesh = dynamic_cast<const TriangleMesh *>(obj.get());
        if (!mesh) throw std::runtime_error("Flatten: only triangle meshes are supported");
        const BSDF *bsdf = mesh->bsdf.get();
        auto mi = matIndex.find(bsdf);
        if (mi == matIndex.end()) {
            std::vector<Float> ser(GetMaxBSDFSerializedSize(), Float(0));
            bsdf->Serialize(Vector2(Float(0), Float(0)), ser.data());
            float rec[10] = {0}, mat[16];
            for (int k = 0; k < 10 && k < int(ser.size()); k++) rec[k] = float(ser[k]);
            bool twoSided = false;
            if (const Lambertian *b = dynamic_cast<const Lambertian *>(bsdf)) twoSided = b->twoSided;
            if (const Phong *b = dynamic_cast<const Phong *>(bsdf)) twoSided = b->twoSided;
            if (drmlt_material_from_serialized(rec, twoSided ? 1 : 0, mat) != DRMLT_OK)
                throw std::runtime_error("Flatten: unsupported BSDF");
            mi = matIndex.emplace(bsdf, int(matIndex.size())).first;
            f.materials.insert(f.materials.end(), mat, mat + 16);
        }
        int li = -1;
        if (mesh->areaLight) {
            const auto it = lightIndex.find(mesh->areaLight);
            if (it == lightIndex.end()) throw std::runtime_error("Flatten: an area light missing from scene.lights");
            li = it->second;
            f.lightTris[2 * li] = int32_t(f.triMat.size());
            f.lightTris[2 * li + 1] = int32_t(mesh->data->indices.size());
            for (int k = 0; k < 3; k++) f.lightEmit[4 * li + k] = float(mesh->areaLight->emission[k]);
            f.lightEmit[4 * li + 3] = float(mesh->areaLight->samplingWeight);
        }
        const TriMeshData &d = *mesh->data;
        for (const TriIndex &t : d.indices) {
            Vector3 p[3], n[3];
            for (int k = 0; k < 3; k++) p[k] = d.position0[t.index[k]];
            const Vector3 gn = Normalize(Cross(Vector3(p[1] - p[0]), Vector3(p[2] - p[0])));
            for (int k = 0; k < 3; k++) n[k] = d.normal0.empty() ? gn : d.normal0[t.index[k]];
            for (int k = 0; k < 3; k++)
                for (int a = 0; a < 3; a++) {
                    f.triP.push_back(float(p[k][a]));
                    f.triN.push_back(float(n[k][a]));
                }
            f.triMat.push_back(mi->second);
            f.triLight.push_back(li);
        }
    }
    for (size_t k = 0; k < lightIndex.size(); k++)
        if (f.lightTris[2 * k] < 0) throw std::runtime_error("Flatten: an area light on no mesh of scene.objects");
    const Camera &cam = *scene.camera;
    const Matrix4x4 m = Interpolate(cam.camToWorld, Float(0));
    const int W = GetPixelWidth(&cam), H = GetPixelHeight(&cam);
    const double fov = 2.0 * std::atan(double(W) / (2.0 * double(cam.dist))) * 180.0 / M_PI;
    for (int a = 0; a < 3; a++) {
        f.camera[a] = float(m(a, 3));                  // position
        f.camera[3 + a] = float(m(a, 3) + m(a, 2));    // a point along the view direction
        f.camera[6 + a] = float(m(a, 1));              // up
    }
    f.camera[9] = float(fov);
    f.camera[10] = float(W);
    f.camera[11] = float(H);
    f.camera[12] = float(cam.nearClip);
    f.camera[13] = float(cam.farClip);
    return f;
}

}  // namespace drmlt_ref
