// scene.hpp - host scene container with the reference's oka::Scene interface
// (include/scene/scene.h:36-423, src/scene/scene.cpp).  It owns the global
// vertex and index buffers, meshes, instances and lights and produces the
// flat GPU records (StkVertex, StkInstance, StkLight) the C ABI consumes.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/stk/stk_types.h"
#include "geom.hpp"

namespace stk {

struct Mat4 {
    float m[16]; // column-major
    static Mat4 identity();
    static Mat4 translate(float x, float y, float z);
    static Mat4 scale(float x, float y, float z);
    static Mat4 rotateX(float radians);
    Mat4 operator*(const Mat4& o) const;
};

// Inverse computed in double precision (rounded to float once).
Mat4 inverse(const Mat4& a);
Mat4 transpose(const Mat4& a);

struct Mesh {
    uint32_t index = 0; // first index in the global index buffer
    uint32_t count = 0; // number of indices
};

struct Instance {
    uint32_t meshId = 0;
    int32_t materialId = -1;
    int32_t lightId = -1;
    Mat4 transform = Mat4::identity();
};

// UniformLightDesc (include/scene/scene.h:131-150), rect/disc/sphere.
struct LightDesc {
    int32_t type = 0; // 0 rect, 1 disc, 2 sphere
    Mat4 xform = Mat4::identity();
    float color[3] = {1.0f, 1.0f, 1.0f};
    float intensity = 1.0f;
    float width = 1.0f, height = 1.0f, radius = 1.0f;
};

// 10:10:10 normal packing of the reference (src/hdStrelka/RenderPass.cpp:50-57).
uint32_t packNormal(F3 n);
// 16:16 UV packing (src/hdStrelka/RenderPass.cpp:59-64).
uint32_t packUV(float u, float v);

class Scene {
public:
    // Scene::createMesh (src/scene/scene.cpp:20-49): appends vertices and
    // rebases indices into the global buffers.
    uint32_t createMesh(const std::vector<StkVertex>& vb, const std::vector<uint32_t>& ib);
    // Scene::createInstance (src/scene/scene.cpp:51-84).
    uint32_t createInstance(uint32_t meshId, int32_t materialId, const Mat4& transform, int32_t lightId = -1);
    // Scene::createLight / updateLight (src/scene/scene.cpp:418-511).
    uint32_t createLight(const LightDesc& desc);
    // Scene::updateLight (src/scene/scene.cpp:452-511): the light record and
    // the transform of its mesh instance; the light type cannot change.
    bool updateLight(uint32_t lightId, const LightDesc& desc);
    // Scene::updateInstanceTransform (include/scene/scene.h:389).
    bool updateInstanceTransform(uint32_t instId, const Mat4& transform);
    uint32_t addMaterial(const StkMaterial& m);

    // createRectLightMesh (src/scene/scene.cpp:117-140)
    uint32_t createRectLightMesh();
    // createDiscLightMesh (src/scene/scene.cpp:318-360)
    uint32_t createDiscLightMesh();
    // createSphereLightMesh (src/scene/scene.cpp:210-316): icosahedron plus
    // midpoint subdivision.  refCompat reproduces the reference's two faces
    // (10,2,1) and (10,11,3); otherwise the closed faces (10,6,1), (10,11,6).
    uint32_t createIcosphereMesh(int subdivisions, bool refCompat);

    const std::vector<StkVertex>& getVertices() const { return vertices_; }
    const std::vector<uint32_t>& getIndices() const { return indices_; }
    const std::vector<Mesh>& getMeshes() const { return meshes_; }
    const std::vector<StkLight>& getLights() const { return lights_; }
    const std::vector<StkMaterial>& getMaterials() const { return materials_; }
    // createInstanceBuffer (src/render/ptrender.cpp:560-601).
    std::vector<StkInstance> getInstanceConstants() const;
    size_t numInstances() const { return instances_.size(); }
    uint32_t lightInstance(uint32_t lightId) const { return lightInstance_[lightId]; }

private:
    std::vector<StkVertex> vertices_;
    std::vector<uint32_t> indices_;
    std::vector<Mesh> meshes_;
    std::vector<Instance> instances_;
    std::vector<StkLight> lights_;
    std::vector<LightDesc> lightDescs_;
    std::vector<uint32_t> lightInstance_; // Scene::mLightIdToInstanceId
    std::vector<StkMaterial> materials_;
    int32_t rectLightMesh_ = -1;
};

// Icosphere vertex/index data (radius 1, outward normals packed).
void makeIcosphere(int subdivisions, bool refCompat, std::vector<StkVertex>& vb, std::vector<uint32_t>& ib);

// Camera matrices of oka::Camera (src/scene/camera.cpp:61-131): reverse-Z
// perspective with its hand-written inverse and a look-at view.
void cameraMatrices(F3 eye, F3 target, F3 up, float vfovDeg, uint32_t width, uint32_t height, float znear,
                    float zfar, StkPathTracerParam& p);

} // namespace stk
