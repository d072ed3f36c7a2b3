// ptrender.hpp - header-only C++ facade over the C ABI (stk.h) with the method
// names of the reference renderer, so code written against oka::PtRender
// (include/render/ptrender.h:40-283) and oka::ResourceManager
// (include/resourcemanager/resourcemanager.h:31-66) keeps its shape.
//
//   reference                              facade
//   PtRender::init / cleanup               stk::PtRender::init / cleanup
//   PtRender::setScene(Scene*)             setScene(const Scene&)       (the container of scene.hpp)
//                                          setScene(const SceneData&)   (arrays Scene::get* returns)
//   PtRender::setActiveCameraIndex(i)      setActiveCameraIndex(i)
//   PtRender::drawFrame(Image*)            drawFrame(Image&)            (RGBA32F or RGBA8, pitch, host or device)
//   PtRender::reloadPt()                   reloadPt()                   (restarts accumulation)
//   SettingsManager "render/pt/..." keys   setSetting(key, value)
//   ResourceManager::createBuffer/...      stk::ResourceManager
//
// Errors: the reference asserts or returns bool/Result; the facade returns
// stk_status from every call that can fail and keeps the library's message in
// lastError().  Nothing here throws.
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "scene.hpp"
#include "stk.h"

namespace stk {

// The arrays oka::Scene hands to the renderer (Scene::getVertices/getIndices/
// getInstances/getLights, include/scene/scene.h:196-274) in their GPU layouts,
// plus the material table and the cameras (view matrices and viewport in
// StkPathTracerParam form, as Camera::setPerspective produces them).
struct SceneData {
    const StkVertex* vertices = nullptr;
    uint32_t numVertices = 0;
    const uint32_t* indices = nullptr;
    uint32_t numIndices = 0;
    const StkInstance* instances = nullptr;
    uint32_t numInstances = 0;
    const StkLight* lights = nullptr;
    uint32_t numLights = 0;
    const StkMaterial* materials = nullptr;
    uint32_t numMaterials = 0;
    std::vector<StkPathTracerParam> cameras;
};

// The image a frame is written to: RGBA32F or RGBA8 pixels, a row pitch (0 =
// packed), in host memory or behind a HIP device (or mapped host) pointer.
struct Image {
    void* data = nullptr;
    uint32_t width = 0, height = 0;
    bool onDevice = false;
    uint32_t format = STK_IMAGE_RGBA32F;
    uint32_t pitchBytes = 0;
};

class PtRender {
public:
    explicit PtRender(int hipDevice = 0) : mDevice(hipDevice) {}
    ~PtRender() { cleanup(); }
    PtRender(const PtRender&) = delete;
    PtRender& operator=(const PtRender&) = delete;

    stk_status init()
    {
        if (mCtx) return STK_OK;
        const stk_status st = stk_create(mDevice, &mCtx);
        if (st != STK_OK) mError = "stk_create failed";
        return st;
    }

    void cleanup()
    {
        if (mCtx) stk_destroy(mCtx);
        mCtx = nullptr;
    }

    // Shader reload has no counterpart (the kernels are compiled code); it
    // restarts the accumulation like a reload does in the reference.
    stk_status reloadPt() { return mCtx ? check(stk_reset_accumulation(mCtx)) : STK_E_STATE; }

    // Uploads the scene and builds its BVH (the reference reads the scene
    // once, at the first frame: src/render/ptrender.cpp:656-667).
    stk_status setScene(const SceneData& s)
    {
        if (!mCtx) return STK_E_STATE;
        stk_status st;
        if ((st = check(stk_set_geometry(mCtx, s.vertices, s.numVertices, s.indices, s.numIndices, s.instances,
                                         s.numInstances))) != STK_OK)
            return st;
        if ((st = check(stk_set_lights(mCtx, s.lights, s.numLights))) != STK_OK) return st;
        if ((st = check(stk_set_materials(mCtx, s.materials, s.numMaterials))) != STK_OK) return st;
        if ((st = check(stk_build_bvh(mCtx, nullptr))) != STK_OK) return st;
        mCameras = s.cameras;
        mCamera = 0;
        mFrame = 0;
        return check(stk_reset_accumulation(mCtx));
    }

    // The scene container of scene.hpp: its buffers go to the device, the
    // BVH is built over them, its cameras become the selectable cameras and
    // its sky the miss radiance.
    stk_status setScene(const Scene& s)
    {
        if (!mCtx) return STK_E_STATE;
        if (!s.valid()) return STK_E_INVALID;
        stk_status st;
        if ((st = check(stk_set_scene(mCtx, s.handle(), nullptr))) != STK_OK) return st;
        mCameras = s.getCameras();
        s.getSky(mSky);
        mCamera = 0;
        mFrame = 0;
        return check(stk_reset_accumulation(mCtx));
    }

    stk_status setActiveCameraIndex(uint32_t index)
    {
        if (index >= mCameras.size()) return STK_E_INVALID;
        if (index != mCamera && mCtx) stk_reset_accumulation(mCtx);
        mCamera = index;
        return STK_OK;
    }
    uint32_t getActiveCameraIndex() const { return mCamera; }

    // The SettingsManager keys drawFrame reads (src/render/ptrender.cpp:671-760
    // and src/hdRunner/main.cpp:414-428).  Unknown keys are rejected.
    stk_status setSetting(const char* key, uint32_t value)
    {
        const std::string k = key ? key : "";
        if (k == "render/pt/spp")
            mSpp = value;
        else if (k == "render/pt/depth")
            mDepth = value;
        else if (k == "render/pt/enableAcc")
            mAccumulate = value != 0;
        else if (k == "render/pt/enableTonemap")
            mTonemap = value != 0;
        else if (k == "render/pt/stratifiedSamplingType")
            mSampling = value;
        else if (k == "render/pt/tonemapperType")
            mTonemapper = value;
        else
            return STK_E_INVALID;
        return STK_OK;
    }
    // Constant miss radiance (the reference adds none; zero reproduces it).
    void setSky(float r, float g, float b)
    {
        mSky[0] = r;
        mSky[1] = g;
        mSky[2] = b;
    }

    // One frame: spp samples per pixel of the active camera, accumulated
    // progressively and tonemapped into `result` (its size is the camera's
    // viewport).
    stk_status drawFrame(Image& result)
    {
        if (!mCtx || mCameras.empty()) return STK_E_STATE;
        StkRenderParams p;
        std::memset(&p, 0, sizeof(p));
        p.pt = mCameras[mCamera];
        if ((uint32_t)p.pt.dimension[0] != result.width || (uint32_t)p.pt.dimension[1] != result.height)
            return STK_E_INVALID;
        p.pt.spp = (int32_t)mSpp;
        p.pt.maxDepth = mDepth;
        p.pt.frameNumber = mFrame++;
        p.pt.stratifiedSamplingType = mSampling;
        p.rrStartDepth = 3;
        p.refCompat = 1;
        p.tonemapperType = mTonemapper;
        p.enableAccumulation = mAccumulate ? 1u : 0u;
        p.enableTonemap = mTonemap ? 1u : 0u;
        for (int i = 0; i < 3; ++i) p.skyColor[i] = mSky[i];
        if (result.format == STK_IMAGE_RGBA32F && result.pitchBytes == 0)
            return check(stk_render(mCtx, &p, static_cast<float*>(result.data), result.onDevice ? 1 : 0));
        StkImageDesc d;
        std::memset(&d, 0, sizeof(d));
        d.data = result.data;
        d.width = result.width;
        d.height = result.height;
        d.pitchBytes = result.pitchBytes;
        d.format = result.format;
        d.onDevice = result.onDevice ? 1 : 0;
        return check(stk_render_image(mCtx, &p, &d));
    }

    stk_ctx* context() const { return mCtx; }
    const std::string& lastError() const { return mError; }

private:
    stk_status check(stk_status st)
    {
        if (st != STK_OK) mError = mCtx ? stk_last_error(mCtx) : "no context";
        return st;
    }

    int mDevice;
    stk_ctx* mCtx = nullptr;
    std::string mError;
    std::vector<StkPathTracerParam> mCameras;
    uint32_t mCamera = 0, mFrame = 0;
    uint32_t mSpp = 1, mDepth = 8, mSampling = 1, mTonemapper = 1;
    bool mAccumulate = true, mTonemap = true;
    float mSky[3] = {0.0f, 0.0f, 0.0f};
};

// Device buffers with oka::ResourceManager's method names, backed by the
// context's HIP allocations and copies.
class ResourceManager {
public:
    enum class MemoryUsage { eGpuOnly, eCpuToGpu }; // device memory, or mapped pinned host memory

    explicit ResourceManager(stk_ctx* ctx) : mCtx(ctx) {}
    void* createBuffer(size_t bytes, MemoryUsage usage = MemoryUsage::eGpuOnly)
    {
        void* p = nullptr;
        if (usage == MemoryUsage::eCpuToGpu) {
            if (stk_host_alloc(mCtx, bytes, &p) != STK_OK) return nullptr;
            mMapped.push_back(p);
            return p;
        }
        return stk_device_alloc(mCtx, bytes, &p) == STK_OK ? p : nullptr;
    }
    // Host address of a host-visible buffer (the same pointer: the memory is
    // mapped into the device's address space); nullptr for device memory.
    void* getMappedMemory(void* buffer) const
    {
        for (void* m : mMapped)
            if (m == buffer) return buffer;
        return nullptr;
    }
    void destroyBuffer(void* p)
    {
        for (size_t i = 0; i < mMapped.size(); ++i)
            if (mMapped[i] == p) {
                mMapped.erase(mMapped.begin() + (long)i);
                stk_host_free(mCtx, p);
                return;
            }
        stk_device_free(mCtx, p);
    }
    stk_status copyToBuffer(void* dst, const void* src, size_t bytes) { return stk_copy_to_device(mCtx, dst, src, bytes); }
    stk_status copyFromBuffer(void* dst, const void* src, size_t bytes)
    {
        return stk_copy_to_host(mCtx, dst, src, bytes);
    }
    stk_status copyBuffer(void* dst, const void* src, size_t bytes) { return stk_device_copy(mCtx, dst, src, bytes); }
    // An RGBA32F image in device memory.
    Image createImage(uint32_t width, uint32_t height)
    {
        Image im;
        im.data = createBuffer((size_t)width * height * 16);
        im.width = im.data ? width : 0;
        im.height = im.data ? height : 0;
        im.onDevice = true;
        return im;
    }

private:
    stk_ctx* mCtx;
    std::vector<void*> mMapped;
};

} // namespace stk
