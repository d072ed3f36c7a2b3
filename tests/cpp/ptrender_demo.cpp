// Renders the small seeded scene (config 6) through the PtRender-shaped C++
// facade and writes the RGBA32F frame to the file named by argv[1].
// Exit codes: 0 rendered, 3 no usable HIP device, 1 any other failure.
#include <cstdio>

#include "stk/ptrender.hpp"
#include "stk/stk_scene.h"

int main(int argc, char** argv)
{
    std::printf("%s\n", stk_version());
    stk_scene* gen = nullptr;
    if (stk_scene_generate(6, 0, 0, &gen) != STK_OK) return 1;
    StkSceneArrays a;
    stk_scene_get(gen, &a);
    stk::SceneData scene;
    scene.vertices = a.vb;
    scene.numVertices = a.numVertices;
    scene.indices = a.ib;
    scene.numIndices = a.numIndices;
    scene.instances = a.inst;
    scene.numInstances = a.numInstances;
    scene.lights = a.lights;
    scene.numLights = a.numLights;
    scene.materials = a.mats;
    scene.numMaterials = a.numMaterials;
    scene.cameras.push_back(a.camera);

    stk::PtRender render(0);
    const stk_status st = render.init();
    if (st == STK_E_DEVICE) {
        std::printf("no device\n");
        stk_scene_free(gen);
        return 3;
    }
    if (st != STK_OK || render.setScene(scene) != STK_OK) return 1;
    render.setSetting("render/pt/spp", 2);
    render.setSetting("render/pt/depth", 8);
    render.setSetting("render/pt/stratifiedSamplingType", 1);
    render.setSky(a.sky[0], a.sky[1], a.sky[2]);
    if (render.setSetting("render/pt/unknown", 1) != STK_E_INVALID) return 1;
    std::vector<float> pixels((size_t)a.camera.dimension[0] * a.camera.dimension[1] * 4);
    stk::Image img;
    img.data = pixels.data();
    img.width = (uint32_t)a.camera.dimension[0];
    img.height = (uint32_t)a.camera.dimension[1];
    for (int f = 0; f < 3; ++f)
        if (render.drawFrame(img) != STK_OK) {
            std::printf("drawFrame: %s\n", render.lastError().c_str());
            return 1;
        }
    // an RGBA8 frame with a padded row pitch into mapped host memory
    stk::ResourceManager rm(render.context());
    const uint32_t pitch = img.width * 4 + 64;
    void* mapped = rm.createBuffer((size_t)pitch * img.height, stk::ResourceManager::MemoryUsage::eCpuToGpu);
    if (!mapped || rm.getMappedMemory(mapped) != mapped) return 1;
    stk::Image img8;
    img8.data = mapped;
    img8.width = img.width;
    img8.height = img.height;
    img8.format = STK_IMAGE_RGBA8;
    img8.pitchBytes = pitch;
    if (render.drawFrame(img8) != STK_OK) {
        std::printf("drawFrame rgba8: %s\n", render.lastError().c_str());
        return 1;
    }
    const unsigned char* px = static_cast<const unsigned char*>(mapped);
    if (px[3] != 255 || px[(size_t)pitch * (img.height - 1) + 3] != 255) return 1; // alpha of rows 0 and H-1
    rm.destroyBuffer(mapped);
    stk_scene_free(gen);
    FILE* out = argc > 1 ? std::fopen(argv[1], "wb") : nullptr;
    if (!out) return 1;
    std::fwrite(pixels.data(), sizeof(float), pixels.size(), out);
    std::fclose(out);
    std::printf("rendered %ux%u\n", img.width, img.height);
    return 0;
}
