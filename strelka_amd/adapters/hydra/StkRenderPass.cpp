// StkRenderPass.cpp - a Hydra render pass that feeds the reference's Hydra
// delegate (src/hdStrelka/) into this library instead of oka::PtRender: it
// bakes the render index's meshes, instancers, materials and lights into the
// scene container of include/stk/scene.hpp and draws frames through
// include/stk/ptrender.hpp.  It follows HdStrelkaRenderPass
// (src/hdStrelka/RenderPass.cpp:24-367) call for call:
//
//   _BakeMeshInstance   RenderPass.cpp:66-126   createMesh + createInstance per mesh instance
//   _BakeMeshes         RenderPass.cpp:128-227  material table, instancer transforms
//   _Execute            RenderPass.cpp:229-367  one-time bake, lights, setScene, drawFrame into the AOV
//
// NOT COMPILED IN THIS REPOSITORY.  It needs the OpenUSD (pxr) headers and
// libraries and the delegate's own prim classes (Mesh.h, Instancer.h,
// Material.h, Light.h, Camera.h, RenderBuffer.h of src/hdStrelka/), none of
// which this image has; strelka_amd/build.py does not build it.  Without
// STK_WITH_USD defined the file is empty.  A maintainer drops it into
// src/hdStrelka/, defines STK_WITH_USD, links libstrelka_amd.so and
// constructs HdStkRenderPass where RenderDelegate.cpp constructs
// HdStrelkaRenderPass (the camera prims then add their cameras to the
// stk::Scene through addCamera/updateCamera, as Camera.cpp:23 and :55 do).
//
// Materials: the reference compiles MDL/MaterialX to HLSL; this library
// shades the Lambert/GGX record of include/stk/stk_types.h.  A mesh display
// colour becomes a Lambert albedo (the reference's "_color" material); a
// network's parameters map by name (diffuse_color / base_color / tint ->
// albedo, roughness / specular_roughness -> GGX roughness, metallic > 0.5 ->
// GGX); anything else renders with the default material.
#ifdef STK_WITH_USD

#include <cstring>
#include <string>
#include <vector>

#include <pxr/base/gf/matrix4d.h>
#include <pxr/imaging/hd/renderPass.h>
#include <pxr/imaging/hd/renderPassState.h>
#include <pxr/imaging/hd/renderIndex.h>
#include <pxr/imaging/hd/rprim.h>

#include "Camera.h"
#include "Instancer.h"
#include "Light.h"
#include "Material.h"
#include "Mesh.h"
#include "RenderBuffer.h"

#include "stk/ptrender.hpp"
#include "stk/scene.hpp"

PXR_NAMESPACE_OPEN_SCOPE

namespace {

void toColumnMajor(const GfMatrix4d& m, float out[16])
{
    // GfMatrix4d is row-vector (translation in row 3); read as glm's
    // column-major array it is the column-vector matrix the library takes,
    // the same element copy RenderPass.cpp:113-120 makes into glm::float4x4
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[i * 4 + j] = (float)m[i][j];
}

StkMaterial materialFromParams(const std::vector<oka::MaterialManager::Param>& params)
{
    StkMaterial m;
    std::memset(&m, 0, sizeof(m));
    m.type = STK_MAT_LAMBERT;
    m.roughness = 1.0f;
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.8f;
    m.albedo[3] = 1.0f;
    for (const auto& p : params) {
        using T = oka::MaterialManager::Param::Type;
        if (p.type == T::eFloat3 && p.value.size() >= 12 &&
            (p.name == "diffuse_color" || p.name == "base_color" || p.name == "tint"))
            std::memcpy(m.albedo, p.value.data(), 12);
        else if (p.type == T::eFloat && p.value.size() >= 4 &&
                 (p.name == "roughness" || p.name == "specular_roughness")) {
            std::memcpy(&m.roughness, p.value.data(), 4);
            m.type = STK_MAT_GGX;
        } else if (p.type == T::eFloat && p.value.size() >= 4 && p.name == "metallic") {
            float metallic = 0.0f;
            std::memcpy(&metallic, p.value.data(), 4);
            if (metallic > 0.5f && m.type != STK_MAT_GLASS) m.type = STK_MAT_GGX;
        } else if (p.type == T::eFloat3 && p.value.size() >= 12 &&
                   (p.name == "emissive_color" || p.name == "emission_color")) {
            std::memcpy(m.emission, p.value.data(), 12);
        } else if (p.type == T::eFloat && p.value.size() >= 4 && p.name == "ior") {
            std::memcpy(&m.ior, p.value.data(), 4);
        } else if (p.type == T::eFloat && p.value.size() >= 4 &&
                   (p.name == "transmission" || p.name == "transmission_weight")) {
            float t = 0.0f;
            std::memcpy(&t, p.value.data(), 4);
            if (t > 0.5f) m.type = STK_MAT_GLASS;
        }
    }
    // a GGX lobe of (near) zero roughness is the mirror's delta
    if (m.type == STK_MAT_GGX && m.roughness < 0.02f) m.type = STK_MAT_MIRROR;
    return m;
}

StkLightDesc lightFrom(const oka::Scene::UniformLightDesc& d)
{
    // HdStrelkaLight always fills the xform form (Light.cpp:149-160)
    StkLightDesc out;
    std::memset(&out, 0, sizeof(out));
    out.type = d.type;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) out.xform[c * 4 + r] = d.xform[c][r];
    for (int k = 0; k < 3; ++k) out.color[k] = d.color[k];
    out.intensity = d.intensity;
    out.width = d.width;
    out.height = d.height;
    out.radius = d.radius;
    return out;
}

} // namespace

class HdStkRenderPass final : public HdRenderPass {
public:
    HdStkRenderPass(HdRenderIndex* index, const HdRprimCollection& collection, stk::PtRender* renderer,
                    stk::Scene* scene)
        : HdRenderPass(index, collection), mRenderer(renderer), mScene(scene)
    {
    }

    bool IsConverged() const override { return mConverged; }

protected:
    void _Execute(const HdRenderPassStateSharedPtr& state, const TfTokenVector& renderTags) override
    {
        TF_UNUSED(renderTags);
        mConverged = false;
        const auto* camera = dynamic_cast<const HdStrelkaCamera*>(state->GetCamera());
        if (!camera) return;
        const HdRenderPassAovBinding* color = nullptr;
        for (const HdRenderPassAovBinding& b : state->GetAovBindings()) {
            if (b.aovName == HdAovTokens->color)
                color = &b;
            else if (auto* rb = dynamic_cast<HdStrelkaRenderBuffer*>(b.renderBuffer))
                rb->SetConverged(true);
        }
        if (!color) return;
        auto* buffer = dynamic_cast<HdStrelkaRenderBuffer*>(color->renderBuffer);
        if (!buffer) return;
        HdRenderIndex* index = GetRenderIndex();

        if (!mBaked) {
            // the reference bakes once, at the first frame (RenderPass.cpp:294-361)
            mBaked = true;
            bakeMeshes(index);
            for (const TfToken& type : {HdPrimTypeTokens->rectLight, HdPrimTypeTokens->diskLight,
                                        HdPrimTypeTokens->sphereLight}) {
                if (!index->IsSprimTypeSupported(type)) continue;
                for (const SdfPath& path : index->GetSprimSubtree(type, SdfPath::AbsoluteRootPath()))
                    if (auto* light = dynamic_cast<HdStrelkaLight*>(index->GetSprim(type, path))) {
                        const StkLightDesc d = lightFrom(light->getLightDesc());
                        if (mScene->createLight(d) == stk::Scene::kInvalidId)
                            TF_WARN("stk: light %s refused", path.GetText());
                    }
            }
            if (mRenderer->setScene(*mScene) != STK_OK) {
                TF_CODING_ERROR("stk::PtRender::setScene: %s", mRenderer->lastError().c_str());
                return;
            }
            mRenderer->setActiveCameraIndex(camera->GetCameraIndex());
        }

        stk::Image image;
        image.data = buffer->Map();
        image.width = buffer->GetWidth();
        image.height = buffer->GetHeight();
        image.format = buffer->GetFormat() == HdFormatUNorm8Vec4 ? STK_IMAGE_RGBA8 : STK_IMAGE_RGBA32F;
        image.onDevice = false;
        if (mRenderer->drawFrame(image) != STK_OK)
            TF_WARN("stk::PtRender::drawFrame: %s", mRenderer->lastError().c_str());
        buffer->Unmap();
        mConverged = true;
    }

private:
    // _BakeMeshInstance (RenderPass.cpp:66-126)
    void bakeMeshInstance(const HdStrelkaMesh* mesh, const GfMatrix4d& transform, uint32_t material)
    {
        const std::vector<GfVec3f>& points = mesh->GetPoints();
        const std::vector<GfVec3f>& normals = mesh->GetNormals();
        const std::vector<GfVec3i>& faces = mesh->GetFaces();
        const std::vector<GfVec2f>& uvs = mesh->GetUVs();
        if (points.empty() || faces.empty() || normals.size() != points.size()) return;
        std::vector<StkVertex> vb(points.size());
        float center[3] = {0.0f, 0.0f, 0.0f};
        for (size_t j = 0; j < points.size(); ++j) {
            StkVertex& v = vb[j];
            std::memset(&v, 0, sizeof(v));
            for (int k = 0; k < 3; ++k) {
                v.position[k] = points[j][k];
                center[k] += points[j][k];
            }
            v.normal = stk::Scene::packNormal(normals[j][0], normals[j][1], normals[j][2]);
            if (!uvs.empty()) v.uv = stk::Scene::packUV(uvs[j][0], 1.0f - uvs[j][1]); // flipped v, as the reference
        }
        for (float& c : center) c /= (float)points.size();
        std::vector<uint32_t> ib(faces.size() * 3);
        for (size_t j = 0; j < faces.size(); ++j)
            for (int k = 0; k < 3; ++k) ib[j * 3 + k] = (uint32_t)faces[j][k];
        float xf[16];
        toColumnMajor(transform, xf);
        const uint32_t meshId = mScene->createMesh(vb, ib);
        if (meshId == stk::Scene::kInvalidId || mScene->createInstance(meshId, material, xf, center) ==
                                                    stk::Scene::kInvalidId)
            TF_WARN("stk: mesh %s refused", mesh->GetId().GetText());
    }

    // _BakeMeshes (RenderPass.cpp:128-227)
    void bakeMeshes(HdRenderIndex* index)
    {
        TfHashMap<SdfPath, uint32_t, SdfPath::Hash> materials;
        for (const SdfPath& id : index->GetRprimIds()) {
            const auto* mesh = dynamic_cast<const HdStrelkaMesh*>(index->GetRprim(id));
            if (!mesh) continue;
            VtMatrix4dArray transforms(1, GfMatrix4d(1.0));
            const SdfPath& instancerId = mesh->GetInstancerId();
            if (!instancerId.IsEmpty())
                if (auto* inst = dynamic_cast<HdStrelkaInstancer*>(index->GetInstancer(instancerId)))
                    transforms = inst->ComputeInstanceTransforms(mesh->GetId());

            uint32_t material = stk::Scene::kInvalidId; // the default material
            const SdfPath& materialId = mesh->GetMaterialId();
            if (mesh->HasColor()) {
                StkMaterial m = materialFromParams({});
                const GfVec3f c = mesh->GetColor();
                for (int k = 0; k < 3; ++k) m.albedo[k] = c[k];
                material = mScene->addMaterial(m);
            } else if (!materialId.IsEmpty()) {
                auto it = materials.find(materialId);
                if (it != materials.end()) {
                    material = it->second;
                } else if (auto* sprim = dynamic_cast<HdStrelkaMaterial*>(
                               index->GetSprim(HdPrimTypeTokens->material, materialId))) {
                    material = mScene->addMaterial(materialFromParams(sprim->getParams()));
                    materials[materialId] = material;
                }
            }
            const GfMatrix4d& prototype = mesh->GetPrototypeTransform();
            for (const GfMatrix4d& t : transforms) bakeMeshInstance(mesh, prototype * t, material);
        }
    }

    stk::PtRender* mRenderer;
    stk::Scene* mScene;
    bool mBaked = false;
    bool mConverged = false;
};

PXR_NAMESPACE_CLOSE_SCOPE

#endif // STK_WITH_USD
