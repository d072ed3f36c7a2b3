// scene.hpp - header-only C++ scene container with the method names of the
// reference's oka::Scene (include/scene/scene.h:36-423), over the stk_scene
// C ABI (stk_scene.h).  Producers written against oka::Scene - the Hydra
// render pass bakes meshes, instances, materials and lights through it
// (src/hdStrelka/RenderPass.cpp:65-220) and the camera prims add cameras
// (src/hdStrelka/Camera.cpp:23) - keep their shape:
//
//   reference                                   facade
//   Scene::createMesh(vb, ib)                   createMesh(vb, ib)
//   Scene::createInstance(mesh, mat, xf, mc)    createInstance(mesh, mat, xf, mc)   xf: column-major float[16]
//   Scene::addMaterial(MaterialDescription)     addMaterial(StkMaterial)            the GPU material record
//   Scene::createLight(UniformLightDesc)        createLight(StkLightDesc)           useXform form
//   Scene::updateLight(id, desc)                updateLight(id, desc)
//   Scene::updateInstanceTransform(id, xf)      updateInstanceTransform(id, xf)
//   Scene::addCamera / getCamera / ...          addCamera / getCamera / updateCamera / getCameraCount
//   Scene::getVertices / getIndices / ...       getVertices / getIndices / getLights / getMaterials / getInstances
//   Scene::mLightIdToInstanceId[id]             getLightInstance(id)
//
// Like the reference, the create/add calls return an id and (uint32_t)-1 when
// they fail (the reference asserts on that value); the get* views point into
// the container and are invalidated by the next create/add/update call.
// Nothing here throws.  The class lives in an inline namespace so that its
// inline members never share a symbol with the library's internal container.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include "stk.h"
#include "stk_scene.h"

namespace stk {
inline namespace v1 {

// A read-only array view (pointer and count).
template <class T>
struct View {
    const T* ptr = nullptr;
    size_t count = 0;
    const T* begin() const { return ptr; }
    const T* end() const { return ptr + count; }
    const T* data() const { return ptr; }
    size_t size() const { return count; }
    bool empty() const { return count == 0; }
    const T& operator[](size_t i) const { return ptr[i]; }
};

class Scene {
public:
    using Vertex = StkVertex;
    using Light = StkLight;
    using Material = StkMaterial;
    using Instance = StkInstance;
    using UniformLightDesc = StkLightDesc;
    using Camera = StkPathTracerParam;
    static constexpr uint32_t kInvalidId = 0xFFFFFFFFu;

    Scene() { stk_scene_create(&mHandle); }
    // One of the seeded workloads of stk_scene_generate.
    static Scene generate(uint32_t config, uint32_t variant = 0, uint32_t flags = 0)
    {
        Scene s(nullptr);
        stk_scene_generate(config, variant, flags, &s.mHandle);
        return s;
    }
    ~Scene()
    {
        if (mHandle) stk_scene_free(mHandle);
    }
    Scene(Scene&& o) noexcept : mHandle(o.mHandle) { o.mHandle = nullptr; }
    Scene& operator=(Scene&& o) noexcept
    {
        if (this != &o) {
            if (mHandle) stk_scene_free(mHandle);
            mHandle = o.mHandle;
            o.mHandle = nullptr;
        }
        return *this;
    }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    // false when the container could not be created (out of memory, or an
    // unknown generator config).
    bool valid() const { return mHandle != nullptr; }
    const stk_scene* handle() const { return mHandle; }
    stk_scene* handle() { return mHandle; }

    uint32_t createMesh(const std::vector<Vertex>& vb, const std::vector<uint32_t>& ib)
    {
        uint32_t id = kInvalidId;
        if (stk_scene_create_mesh(mHandle, vb.data(), (uint32_t)vb.size(), ib.data(), (uint32_t)ib.size(), &id) !=
            STK_OK)
            return kInvalidId;
        return id;
    }
    // materialId (uint32_t)-1 = the default material; massCenter is accepted
    // for the reference's signature and not stored (its renderer never reads it).
    uint32_t createInstance(uint32_t meshId, uint32_t materialId, const float transform[16],
                            const float* massCenter = nullptr)
    {
        (void)massCenter;
        uint32_t id = kInvalidId;
        const int32_t mat = materialId == kInvalidId ? -1 : (int32_t)materialId;
        if (materialId != kInvalidId && materialId > 0x7FFFFFFFu) return kInvalidId;
        if (stk_scene_create_instance(mHandle, meshId, mat, transform, &id) != STK_OK) return kInvalidId;
        return id;
    }
    bool updateInstanceTransform(uint32_t instId, const float transform[16])
    {
        return stk_scene_update_instance_transform(mHandle, instId, transform) == STK_OK;
    }
    uint32_t addMaterial(const Material& m)
    {
        uint32_t id = kInvalidId;
        return stk_scene_add_material(mHandle, &m, &id) == STK_OK ? id : kInvalidId;
    }
    uint32_t createLight(const UniformLightDesc& desc)
    {
        uint32_t id = kInvalidId;
        return stk_scene_create_light(mHandle, &desc, &id) == STK_OK ? id : kInvalidId;
    }
    bool updateLight(uint32_t lightId, const UniformLightDesc& desc)
    {
        return stk_scene_update_light(mHandle, lightId, &desc) == STK_OK;
    }
    uint32_t getLightInstance(uint32_t lightId) const
    {
        uint32_t id = kInvalidId;
        return stk_scene_light_instance(mHandle, lightId, &id) == STK_OK ? id : kInvalidId;
    }

    uint32_t addCamera(const Camera& camera)
    {
        uint32_t id = kInvalidId;
        return stk_scene_add_camera(mHandle, &camera, &id) == STK_OK ? id : kInvalidId;
    }
    bool updateCamera(const Camera& camera, uint32_t index)
    {
        return stk_scene_update_camera(mHandle, index, &camera) == STK_OK;
    }
    // The camera at index (zero-initialised when index is out of range).
    Camera getCamera(uint32_t index) const
    {
        Camera c;
        if (stk_scene_get_camera(mHandle, index, &c) != STK_OK) std::memset(&c, 0, sizeof(c));
        return c;
    }
    size_t getCameraCount() const { return stk_scene_camera_count(mHandle); }
    std::vector<Camera> getCameras() const
    {
        std::vector<Camera> out(getCameraCount());
        for (uint32_t i = 0; i < out.size(); ++i) out[i] = getCamera(i);
        return out;
    }
    // oka::Camera's look-at form: eye, target and up vectors, vertical field
    // of view in degrees, viewport, near and far planes.
    static Camera lookAt(const float eye[3], const float target[3], const float up[3], float vfovDeg, uint32_t width,
                         uint32_t height, float znear = 0.01f, float zfar = 100.0f)
    {
        Camera c;
        if (stk_camera_look_at(eye, target, up, vfovDeg, width, height, znear, zfar, &c) != STK_OK)
            std::memset(&c, 0, sizeof(c));
        return c;
    }

    // Constant miss radiance (the reference adds none).
    bool setSky(float r, float g, float b) { return stk_scene_set_sky(mHandle, r, g, b) == STK_OK; }

    View<Vertex> getVertices() const { return arrays().vertices; }
    View<uint32_t> getIndices() const { return arrays().indices; }
    View<Instance> getInstances() const { return arrays().instances; }
    View<Light> getLights() const { return arrays().lights; }
    View<Material> getMaterials() const { return arrays().materials; }
    // The sky of setSky (or of the generated workload), RGB.
    void getSky(float out[3]) const
    {
        StkSceneArrays a;
        std::memset(&a, 0, sizeof(a));
        if (mHandle) stk_scene_get(mHandle, &a);
        for (int k = 0; k < 3; ++k) out[k] = a.sky[k];
    }

    // The 10:10:10 normal and 16:16 UV packings of the vertex record
    // (src/hdStrelka/RenderPass.cpp:50-64).
    static uint32_t packNormal(float x, float y, float z) { return stk_pack_normal(x, y, z); }
    static uint32_t packUV(float u, float v) { return stk_pack_uv(u, v); }

private:
    explicit Scene(std::nullptr_t) {}

    struct Arrays {
        View<Vertex> vertices;
        View<uint32_t> indices;
        View<Instance> instances;
        View<Light> lights;
        View<Material> materials;
    };
    Arrays arrays() const
    {
        Arrays r;
        StkSceneArrays a;
        if (!mHandle || stk_scene_get(mHandle, &a) != STK_OK) return r;
        r.vertices = {a.vb, a.numVertices};
        r.indices = {a.ib, a.numIndices};
        r.instances = {a.inst, a.numInstances};
        r.lights = {a.lights, a.numLights};
        r.materials = {a.mats, a.numMaterials};
        return r;
    }

    stk_scene* mHandle = nullptr;
};

} // namespace v1
} // namespace stk
