/*
 * stk_types.h - byte-exact data contracts shared by the HIP library, the C-ABI
 * callers and the CPU oracle.
 *
 * Every struct mirrors a GPU-visible buffer record of the reference renderer;
 * the offsets are pinned with static asserts so that a buffer produced for the
 * reference can be handed to this library unchanged.
 *
 * Matrices are 16 floats in column-major order (element (row r, col c) lives at
 * m[c*4 + r]), i.e. the memory order of glm::mat4, which is what the reference
 * uploads into its constant and structured buffers.
 */
#ifndef STK_TYPES_H
#define STK_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STK_INVALID_INDEX 0xFFFFFFFFu

/* Vertex record, 32 B.  Mirrors `struct Vertex` (shaders/raytracing.h:17-25)
 * and oka::Scene::Vertex (include/scene/scene.h:54-63). */
typedef struct StkVertex {
    float position[3];
    uint32_t tangent; /* 10:10:10 packed, decoded like a normal */
    uint32_t normal;  /* 10:10:10 packed over [-1,1] (shaders/pack.h:4-13) */
    uint32_t uv;      /* 16:16 packed over [-10,10] (shaders/pack.h:15-23) */
    float pad0;
    float pad1;
} StkVertex;

/* BVH node in the reference's flat skip-link layout, 32 B.  Mirrors
 * `struct BVHNode` (shaders/raytracing.h:27-33, include/render/bvh.h:17-24).
 * Inner node: bounds, instId = 0xFFFFFFFF, nodeOffset = skip (miss) link.
 * Leaf: minBounds[0] holds the primitive id bit-cast to float, instId is the
 * instance id, maxBounds = 0, nodeOffset = skip link.  Nodes are stored in DFS
 * preorder; the hit (descend) link of node i is i+1. */
typedef struct StkBvhNode {
    float minBounds[3];
    int32_t instId;
    float maxBounds[3];
    int32_t nodeOffset;
} StkBvhNode;

/* Per-instance constants, 208 B.  Mirrors `struct InstanceConstants`
 * (shaders/instanceconstants.h:19-28), filled as in
 * src/render/ptrender.cpp:574-587. */
typedef struct StkInstance {
    float objectToWorld[16];
    float worldToObject[16];
    float normalMatrix[16];
    int32_t materialId;
    int32_t indexOffset;
    int32_t indexCount;
    int32_t lightId; /* -1 for non-light instances */
} StkInstance;

/* Light record, 112 B.  Mirrors `struct UniformLight` (shaders/lights.h:10-19)
 * and oka::Scene::Light (include/scene/scene.h:120-129).  type 0 = rect (four
 * world-space corners), 1 = disc, 2 = sphere (points[0].x = radius,
 * points[1] = centre). */
typedef struct StkLight {
    float points[4][4];
    float color[4];
    float normal[4];
    int32_t type;
    float pad0;
    float pad2;
    float pad3;
} StkLight;

/* Ray, 32 B.  Mirrors `struct Ray` (shaders/ray.h:12-16): o.xyz origin,
 * o.w = tmax, d.xyz direction, d.w unused. */
typedef struct StkRay {
    float o[4];
    float d[4];
} StkRay;

/* Hit record, 20 B.  Mirrors `struct Hit` (shaders/raytracing.h:90-96).
 * A miss is reported as instId = primId = 0xFFFFFFFF, t = 1e9. */
typedef struct StkHit {
    float bary[2];
    float t;
    uint32_t instId;
    uint32_t primId;
} StkHit;

/* Per-frame path tracer constants, 336 B.  Mirrors `struct PathTracerParam`
 * (shaders/pathtracerparam.h:20-41), filled as in
 * src/render/ptrender.cpp:755-774. */
typedef struct StkPathTracerParam {
    float viewToWorld[16];
    float clipToView[16];
    float viewToClip[16];
    float worldToView[16];
    float camPos[4];
    int32_t dimension[2];
    float invDimension[2];
    uint32_t frameNumber;
    uint32_t maxDepth;
    uint32_t debug;
    uint32_t numLights;
    int32_t len;
    int32_t spp;
    int32_t iteration;
    uint32_t stratifiedSamplingType; /* 0 centre, 1 random, 2 stratified, 3 CMJ */
    float pad0[4];
} StkPathTracerParam;

/* Material record of this library, 64 B.  The reference binds MDL argument
 * blocks here (shaders/pathtracerMdl.hlsl:276) and calls the MDL-compiled
 * scattering and emission functions (:299-372); this library evaluates
 * closed-form models in their place:
 *   STK_MAT_LAMBERT  diffuse reflection (BSDF event diffuse reflection);
 *   STK_MAT_GGX      GGX microfacet reflection, Schlick Fresnel with f0 =
 *                    albedo (glossy reflection);
 *   STK_MAT_MIRROR   ideal specular reflection about the shading normal,
 *                    weight Schlick(albedo, cos) (specular reflection);
 *   STK_MAT_GLASS    smooth dielectric of index `ior` (0 selects 1.5): the
 *                    exact unpolarised Fresnel reflectance F chooses between
 *                    specular reflection (weight 1) and refraction by
 *                    Snell's law (weight albedo, the tint); the side is that
 *                    of the geometric normal (entering where the ray meets
 *                    its front face).
 * Specular events carry no pdf: next-event estimation adds nothing at such a
 * vertex and an emitter the next ray meets counts in full, as after the
 * reference's specularBounce (:229-234, :364).  emission (x the emission
 * texture's rgb) is added, times the path throughput, wherever a ray meets
 * the front face of the surface (the EDF term, :299-304).  Textures are
 * looked up at the hit's texture coordinates (the vertices' packed uv,
 * shaders/pack.h:15-23, interpolated as :261-265): albedo x the albedo
 * texture's rgb, roughness x the roughness texture's red channel, and the
 * normal texture a tangent-space normal map applied as the reference's
 * CalcBumpedNormal (shaders/helper.h:101-115) with the interpolated vertex
 * tangent. */
enum { STK_MAT_LAMBERT = 0, STK_MAT_GGX = 1, STK_MAT_MIRROR = 2, STK_MAT_GLASS = 3 };
typedef struct StkMaterial {
    uint32_t type;
    float roughness;           /* GGX alpha = roughness^2 */
    uint32_t albedoTexture;    /* 0: none; k: texture k - 1 of stk_set_textures */
    uint32_t roughnessTexture; /* 0: none; k: roughness x texture k - 1's red channel */
    float albedo[4];           /* rgb; [3] unused */
    float emission[3];         /* emitted radiance (front face) */
    float ior;                 /* glass: index of refraction (0: 1.5) */
    uint32_t normalTexture;    /* 0: none; k: tangent-space normal map k - 1 (rgb = (n + 1) / 2) */
    uint32_t emissionTexture;  /* 0: none; k: emission x texture k - 1's rgb */
    uint32_t pad[2];
} StkMaterial;

/* A texture of stk_set_textures: width x height RGBA32F texels, row-major
 * (row 0 at v = 0), from texel `offset` of the texel array.  Lookups follow
 * the reference's tex_lookup_float4_2d (shaders/mdl_runtime.hlsl:173-207):
 * the wrap of apply_wrap_and_crop (:93-120; repeat, clamp to the edge
 * texels' centres, mirrored repeat then that clamp, or clip: zero outside
 * [0, 1), clamped inside), the quintic "smootherstep" of the fractional texel
 * position (:125-140), then a bilinear blend at level 0 (the reference's
 * SampleLevel(..., 0), :203-206: no mip levels are read). */
enum { STK_TEX_REPEAT = 0, STK_TEX_CLAMP = 1, STK_TEX_MIRRORED_REPEAT = 2, STK_TEX_CLIP = 3 };
typedef struct StkTexture {
    uint32_t width;
    uint32_t height;
    uint32_t offset; /* first texel (RGBA32F, 16 B) of the texture */
    uint32_t wrap;   /* STK_TEX_*, both axes */
} StkTexture;

/* Pre-transformed world-space triangle, 48 B: {v0, e0, e1} exactly as
 * getTriangle (shaders/raytracing.h:134-153) computes them, plus the ids the
 * hit record needs and the preorder index of the leaf that references it
 * (used for the order-independent tie-break). */
typedef struct StkTriangle {
    float v0[3];
    uint32_t leaf; /* preorder index of the leaf node in the skip-link layout */
    float e0[3];
    uint32_t instId;
    float e1[3];
    uint32_t primId;
} StkTriangle;

/* Packed BVH2 node of the MI355X traversal, 64 B (one aligned line).
 * Children boxes are interleaved per axis so one node is four 16-B loads:
 *   x = {c0.lo.x, c0.hi.x, c1.lo.x, c1.hi.x}, y = ..., z = ...
 * child[i] >= 0 : index of an inner packed node
 * child[i] <  0 : leaf, triangle index = ~child[i] (its box is not tested: the
 *                 reference never tests leaf boxes, shaders/raytracing.h:211-231)
 */
typedef struct StkPackedNode {
    float x[4];
    float y[4];
    float z[4];
    int32_t child[2];
    uint32_t pad[2];
} StkPackedNode;

/* Compressed 8-wide BVH node of the MI355X traversal, 80 B (five 16-B loads).
 * A node holds up to 8 children in slots; slot s of a node is used when bit s
 * of vmask is set, holds an inner node when bit s of imask is set and a leaf
 * (one triangle) otherwise.  Inner children are consecutive nodes from
 * childBase and leaf triangles consecutive from triBase, both in slot order.
 * Child boxes are quantised to 8 bits per plane relative to origin:
 *   lo[a] = origin[a] + qlo[a][s] * 2^(exps[a] - 127)   (float arithmetic)
 *   hi[a] = origin[a] + qhi[a][s] * 2^(exps[a] - 127)
 * rounded outwards, so every decoded box contains its child's box (a leaf's
 * box is its triangle's box widened by a relative 1e-5).  Slots are assigned
 * by the direction of the child's centre from the node's centre (bit a of s
 * set: the positive side of axis a), so visiting slots k ^ octant(ray),
 * k = 0..7, goes roughly front to back. */
typedef struct StkWideNode {
    float origin[3];
    uint8_t exps[3];
    uint8_t imask;
    uint32_t childBase;
    uint32_t triBase;
    uint32_t vmask; /* low 8 bits */
    uint32_t pad;
    uint8_t qlo[3][8];
    uint8_t qhi[3][8];
} StkWideNode;

/* Two-level (TLAS/BLAS) records, for scenes of many instances of few meshes.
 * A BLAS is a compressed 8-wide tree (StkWideNode) over one mesh's triangles
 * in object space; the TLAS is one over the instances' world boxes.  The hit
 * test still runs on the world-space triangle formed exactly as getTriangle
 * (shaders/raytracing.h:134-153) forms it, so hits are those of the
 * reference walk over the same tree expanded into one skip-link array (each
 * instance leaf of the TLAS replaced by its mesh's tree). */
typedef struct StkBlasTriangle {
    float p0[3];
    uint32_t primId; /* triangle index within the mesh (= within each instance) */
    float p1[3];
    uint32_t leaf;   /* preorder index of its leaf in the mesh's binary tree */
    float p2[3];
    uint32_t pad;
} StkBlasTriangle;

typedef struct StkTlasLeaf {
    float lo[3];       /* the instance's world box (holds every triangle as */
    uint32_t instId;   /* getTriangle forms it)                             */
    float hi[3];
    uint32_t blasRoot; /* wide node index of the root of the instance's mesh tree */
    uint32_t key;      /* preorder index of the instance's leaf in the TLAS binary tree */
    uint32_t pad[3];
} StkTlasLeaf;

/* One radiance-cache sample, 64 B: the shading point, its shading normal, the
 * outgoing direction, the roughness and (for training) the radiance target.
 * payload is free for the caller (the renderer stores the path slot there). */
typedef struct StkCacheRecord {
    float position[3];
    float roughness;
    float normal[3];
    uint32_t payload;
    float direction[3];
    float pad0;
    float target[3];
    float pad1;
} StkCacheRecord;

/* Parameters of one render call (progressive; one call adds `spp` samples). */
typedef struct StkRenderParams {
    StkPathTracerParam pt;   /* camera matrices, dimension, spp, maxDepth ... */
    uint32_t rrStartDepth;   /* RR applies when depth > rrStartDepth (reference: 3) */
    uint32_t refCompat;      /* 1 = reproduce the reference estimator exactly */
    uint32_t tonemapperType; /* 0 Reinhard, 1 ACES fitted, 2 ACES film */
    uint32_t enableAccumulation;
    float skyColor[4];       /* constant miss radiance (reference adds 0) */
    uint32_t enableTonemap;
    uint32_t keepSamples;    /* keep per-sample radiance for stk_get_samples */
    /* radiance cache (needs stk_cache_create): 0 off (plain path tracer),
     * 1 query and train, 2 query only (frozen cache), 3 train only, 4 train
     * only and trace the training paths only (warm-up; the image holds only
     * those paths' samples) */
    uint32_t cacheMode;
    /* Paths other than the training paths query the cache at the first
     * surface vertex of depth >= cacheQueryDepth (0 = the camera ray's hit)
     * whose roughness is >= cacheMinRoughness (Lambert counts as 1): the
     * vertex's direct light is sampled as usual (unless cacheQueryDirect),
     * and the radiance scattered from the BSDF-sampled direction on comes
     * from the cache. */
    uint32_t cacheQueryDepth;
    uint32_t cacheTrainPaths; /* training paths per frame (extended to maxDepth, or see cacheTrainSuffix) */
    float cacheMinRoughness;  /* in [0, 1]; 0 queries at cacheQueryDepth */
    uint32_t cacheTrainSteps; /* optimiser steps per frame over the frame's training records (0 = 1, <= 64) */
    /* 0: training paths run to maxDepth.  n > 0 (with cacheMode 1): a training
     * path ends at the first vertex of depth >= cacheQueryDepth + n where the
     * other paths would query, with a query of its own (self-training, Mueller
     * et al. 2021, "Real-time Neural Radiance Caching", section 5): that vertex
     * samples its direct light, the cache stands in for the rest, and the
     * vertex itself is not a training record.  Shorter training paths, a
     * smoother target. */
    uint32_t cacheTrainSuffix;
    /* 0: a query vertex samples its direct light and the cache answers for
     * the rest.  1: the cache answers for all the radiance the vertex
     * reflects, its direct light included (no shadow ray there; the training
     * targets then include each vertex's own light sample), as neural
     * radiance caching queries its cache (Mueller et al. 2021, section 3). */
    uint32_t cacheQueryDirect;
    uint32_t pad[1];
} StkRenderParams;

#ifdef __cplusplus
} /* extern "C" */
#endif

#ifdef __cplusplus
#define STK_STATIC_ASSERT static_assert
#else
#define STK_STATIC_ASSERT _Static_assert
#endif

STK_STATIC_ASSERT(sizeof(StkVertex) == 32, "Vertex must be 32 B");
STK_STATIC_ASSERT(sizeof(StkBvhNode) == 32, "BVHNode must be 32 B");
STK_STATIC_ASSERT(offsetof(StkBvhNode, instId) == 12, "BVHNode.instId offset");
STK_STATIC_ASSERT(offsetof(StkBvhNode, nodeOffset) == 28, "BVHNode.nodeOffset offset");
STK_STATIC_ASSERT(sizeof(StkInstance) == 208, "InstanceConstants must be 208 B");
STK_STATIC_ASSERT(offsetof(StkInstance, materialId) == 192, "InstanceConstants.materialId offset");
STK_STATIC_ASSERT(offsetof(StkInstance, lightId) == 204, "InstanceConstants.lightId offset");
STK_STATIC_ASSERT(sizeof(StkLight) == 112, "UniformLight must be 112 B");
STK_STATIC_ASSERT(offsetof(StkLight, type) == 96, "UniformLight.type offset");
STK_STATIC_ASSERT(sizeof(StkRay) == 32, "Ray must be 32 B");
STK_STATIC_ASSERT(sizeof(StkHit) == 20, "Hit must be 20 B");
STK_STATIC_ASSERT(sizeof(StkPathTracerParam) == 336, "PathTracerParam must be 336 B");
STK_STATIC_ASSERT(offsetof(StkPathTracerParam, dimension) == 272, "PathTracerParam.dimension offset");
STK_STATIC_ASSERT(offsetof(StkPathTracerParam, frameNumber) == 288, "PathTracerParam.frameNumber offset");
STK_STATIC_ASSERT(offsetof(StkPathTracerParam, len) == 304, "PathTracerParam.len offset");
STK_STATIC_ASSERT(sizeof(StkMaterial) == 64, "Material must be 64 B");
STK_STATIC_ASSERT(offsetof(StkMaterial, albedoTexture) == 8, "Material.albedoTexture offset");
STK_STATIC_ASSERT(offsetof(StkMaterial, roughnessTexture) == 12, "Material.roughnessTexture offset");
STK_STATIC_ASSERT(offsetof(StkMaterial, emission) == 32, "Material.emission offset");
STK_STATIC_ASSERT(offsetof(StkMaterial, normalTexture) == 48, "Material.normalTexture offset");
STK_STATIC_ASSERT(sizeof(StkTexture) == 16, "Texture must be 16 B");
STK_STATIC_ASSERT(sizeof(StkTriangle) == 48, "Triangle must be 48 B");
STK_STATIC_ASSERT(sizeof(StkPackedNode) == 64, "PackedNode must be 64 B");
STK_STATIC_ASSERT(sizeof(StkWideNode) == 80, "WideNode must be 80 B");
STK_STATIC_ASSERT(offsetof(StkWideNode, qlo) == 32, "WideNode.qlo offset");
STK_STATIC_ASSERT(sizeof(StkCacheRecord) == 64, "CacheRecord must be 64 B");
STK_STATIC_ASSERT(sizeof(StkBlasTriangle) == 48, "BlasTriangle must be 48 B");
STK_STATIC_ASSERT(sizeof(StkTlasLeaf) == 48, "TlasLeaf must be 48 B");
STK_STATIC_ASSERT(sizeof(StkRenderParams) == 336 + 72, "RenderParams must be 408 B");

#endif /* STK_TYPES_H */
