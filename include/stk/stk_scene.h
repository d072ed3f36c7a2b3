/*
 * stk_scene.h - seeded synthetic workloads (part of libstrelka_amd.so).
 *
 * The reference renderer loads USD stages; it has no synthetic scenes and no
 * benchmark workloads.  These generators build the measured configurations
 * from a PCG32 stream (O'Neill 2014, XSH-RR, multiplier 6364136223846793005)
 * seeded with 0x5354524B00000000 + config, one stream id per purpose
 * (1 geometry, 2 materials, 3 ray jitter).  The meshes go through the same
 * Scene container as user content (createMesh / createInstance / createLight
 * of src/scene/scene.cpp).
 *
 *   config 1: icosphere (3 subdivisions, r = 1, 1,280 triangles) inside an
 *             inward-facing box [-2,2]^3 (12 triangles); camera at (0,0,1.9)
 *             looking down -z, vfov 60, 256x256.
 *   config 2: 512x512-quad fBm heightfield over [-10,10]^2 (524,288
 *             triangles) + 2,000 icospheres of 320 triangles, one 2x2 quad
 *             light at y = 8 facing -y (intensity 100), Lambert/GGX
 *             materials, 1920x1080 camera; 1,164,290 triangles.
 *   config 3: one of the 32 uneven scenes of the batched launch (variant
 *             0..31), built as config 2 with ~N_i triangles, 960x540.
 *   config 5: 20 seeded meshes (10 displaced icospheres of 1,280 triangles,
 *             10 tori of 500-5,000) instanced 10,000 times with uniform random
 *             rotations and translations in [-100,100]^3 (about 20 M
 *             triangles), one Lambert/GGX material per instance, an 80x80 rect
 *             light at y = 130 (radiance 8), sky (0.25, 0.3, 0.4), 3840x2160.
 *             variant n > 0: n instances in a box scaled by cbrt(n / 10,000)
 *             (same density; light and camera scaled alike), 384x216.
 *   config 6: a tiny scene (heightfield 16x16, 8 spheres) for smoke tests.
 *   config 7: the config-1 room closed and lit by a 1x1 ceiling light
 *             (walls albedo 0.7), 128x128: long multi-bounce paths.
 *   config 8: convex furnace: an outward-facing cube [-1,1]^3 (flat faces,
 *             Lambert albedo (0.8, 0.6, 0.4)) under a constant sky
 *             (0.5, 0.7, 1.0), no lights; camera at (3, 2.4, 3.6) looking at
 *             the origin, vfov 40, 96x96.  A path that hits the cube bounces
 *             once and escapes, so every sample is albedo*sky or sky.
 *   config 9: lit floor: a Lambert floor y = 0 over [-20,20]^2 (albedo 0.5)
 *             under a 1x1 rect light centred at y = 2 facing -y (intensity
 *             10, so radiance 10), no sky; camera at (0, 1, 0) looking down,
 *             up -z, vfov 60, 64x64.  A pixel's value is the next-event
 *             estimate at the floor, an integral over the light.
 */
#ifndef STK_SCENE_H
#define STK_SCENE_H

#include "stk.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef STK_SCENE_HANDLE
#define STK_SCENE_HANDLE
typedef struct stk_scene stk_scene;
#endif

typedef struct StkSceneArrays {
    const StkVertex* vb;
    uint32_t numVertices;
    const uint32_t* ib;
    uint32_t numIndices;
    const StkInstance* inst;
    uint32_t numInstances;
    const StkLight* lights;
    uint32_t numLights;
    const StkMaterial* mats;
    uint32_t numMaterials;
    uint32_t numTriangles;
    uint32_t pad;
    StkPathTracerParam camera; /* matrices, dimension; spp = 1, maxDepth = 8 */
    float sky[4];              /* constant sky radiance of the config */
} StkSceneArrays;

/* flags: bit 0 = reproduce the reference icosahedron faces (10,2,1),(10,11,3) */
#define STK_SCENE_REF_ICOSPHERE 0x1u

STK_API stk_status stk_scene_generate(uint32_t config, uint32_t variant, uint32_t flags, stk_scene** out);
STK_API void stk_scene_free(stk_scene* s);
STK_API stk_status stk_scene_get(const stk_scene* s, StkSceneArrays* out);
/* Config-1 style jittered primary rays of the scene camera for every pixel:
 * pixel position = pixel + U[0,1)^2, origin = camera + U[-1e-3,1e-3]^3, PCG
 * stream 3 of `seed`.  rays holds W*H entries, pixel order. */
STK_API stk_status stk_scene_jittered_rays(const stk_scene* s, uint64_t seed, StkRay* rays, uint32_t n);

/* ---------------------------------------------------------------------------
 * Scene building: the calls oka::Scene offers its producers (the Hydra render
 * pass of src/hdStrelka/RenderPass.cpp:65-126 and 128-220, the camera prims of
 * src/hdStrelka/Camera.cpp:23), over the container the generators above fill,
 * so that a scene built through them yields the same arrays (stk_scene_get)
 * and renders the same through stk_set_scene (stk.h).  Matrices are
 * column-major float[16] (glm's layout).  Ids count from 0 in creation order.
 * Every call returns STK_E_INVALID for a null or out-of-range argument and
 * STK_E_OOM when a host allocation fails; calls on one scene must not overlap.
 */

/* UniformLightDesc (include/scene/scene.h:132-150) with useXform = true: the
 * light frame is xform; a rect light spans [-width/2, width/2] x [-height/2,
 * height/2] of its local z = 0 plane and emits along local -z (the normal of
 * its corner order); radiance = color * intensity / area (disc: pi r^2,
 * sphere: 4 pi r^2). */
typedef struct StkLightDesc {
    int32_t type;    /* 0 rect, 1 disc, 2 sphere (only rect lights are sampled by the renderer) */
    float xform[16]; /* light to world */
    float color[3];
    float intensity;
    float width, height; /* rect, > 0 */
    float radius;        /* disc, sphere, > 0 */
} StkLightDesc;

/* An empty scene (no cameras, sky 0). */
STK_API stk_status stk_scene_create(stk_scene** out);
/* Scene::createMesh (src/scene/scene.cpp:20-49): appends the vertices and the
 * indices (local to vb, rebased here) to the global buffers. */
STK_API stk_status stk_scene_create_mesh(stk_scene* s, const StkVertex* vb, uint32_t numVertices,
                                         const uint32_t* ib, uint32_t numIndices, uint32_t* meshId);
/* Scene::createInstance (src/scene/scene.cpp:51-84); materialId -1 = the
 * default material.  The reference's massCenter argument is not stored (its
 * renderer does not read it). */
STK_API stk_status stk_scene_create_instance(stk_scene* s, uint32_t meshId, int32_t materialId,
                                             const float transform[16], uint32_t* instId);
/* Scene::updateInstanceTransform (include/scene/scene.h:389). */
STK_API stk_status stk_scene_update_instance_transform(stk_scene* s, uint32_t instId, const float transform[16]);
/* Scene::addMaterial (src/scene/scene.cpp:86-91) with the GPU material record. */
STK_API stk_status stk_scene_add_material(stk_scene* s, const StkMaterial* m, uint32_t* materialId);
/* Scene::createLight (src/scene/scene.cpp:418-450): the light record and an
 * instance of its emitting mesh (rect, 16-segment disc, icosphere). */
STK_API stk_status stk_scene_create_light(stk_scene* s, const StkLightDesc* desc, uint32_t* lightId);
/* Scene::updateLight (src/scene/scene.cpp:452-511) plus the transform of the
 * light's mesh instance; the type must stay the same. */
STK_API stk_status stk_scene_update_light(stk_scene* s, uint32_t lightId, const StkLightDesc* desc);
/* Scene::mLightIdToInstanceId. */
STK_API stk_status stk_scene_light_instance(const stk_scene* s, uint32_t lightId, uint32_t* instId);
/* Scene::addCamera / updateCamera / getCamera / getCameraCount
 * (include/scene/scene.h:231-264); camera 0 is the one stk_scene_get reports. */
STK_API stk_status stk_scene_add_camera(stk_scene* s, const StkPathTracerParam* camera, uint32_t* cameraId);
STK_API stk_status stk_scene_update_camera(stk_scene* s, uint32_t cameraId, const StkPathTracerParam* camera);
STK_API uint32_t stk_scene_camera_count(const stk_scene* s);
STK_API stk_status stk_scene_get_camera(const stk_scene* s, uint32_t cameraId, StkPathTracerParam* out);
/* The vertex record's 10:10:10 normal packing (components in [-1, 1]) and
 * 16:16 UV packing (components in [-10, 10]) of src/hdStrelka/RenderPass.cpp:
 * 50-64. */
STK_API uint32_t stk_pack_normal(float x, float y, float z);
STK_API uint32_t stk_pack_uv(float u, float v);
/* Constant miss radiance (the reference adds none). */
STK_API stk_status stk_scene_set_sky(stk_scene* s, float r, float g, float b);
/* oka::Camera's view and reverse-Z projection matrices (src/scene/camera.cpp:
 * 61-131) for a look-at camera: viewport width x height, vertical field of
 * view in degrees; spp = 1, maxDepth = 8, the rest zero. */
STK_API stk_status stk_camera_look_at(const float eye[3], const float target[3], const float up[3], float vfovDeg,
                                      uint32_t width, uint32_t height, float znear, float zfar,
                                      StkPathTracerParam* out);

#ifdef __cplusplus
}
#endif
#endif
