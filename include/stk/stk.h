/*
 * stk.h - C ABI of the MI355X path-tracing library (libstrelka_amd.so).
 *
 * This is the drop-in boundary that replaces the Vulkan render passes of the
 * reference renderer (oka::PtRender, include/render/ptrender.h:40-283).  Plain
 * pointers and sizes only; nothing throws across it.  One context owns one HIP
 * stream on one device; calls on a context must be serialised by the caller,
 * different contexts are independent.
 *
 * Reference entry points replaced (file:line in ikryukov/Strelka):
 *   stk_create / stk_destroy     PtRender::init / cleanup / setSharedContext
 *                                (include/render/ptrender.h:47-82,
 *                                 src/render/ptrender.cpp:40-195)
 *   stk_set_geometry             Scene::getVertices/getIndices/getInstances read
 *                                by createVertexBuffer/IndexBuffer/InstanceBuffer
 *                                (src/render/ptrender.cpp:400-601, read once at
 *                                frame 0, :656-667)
 *   stk_set_lights               createLightsBuffer (src/render/ptrender.cpp:436-458)
 *   stk_set_materials            createMdlBuffers (src/render/ptrender.cpp:603-640);
 *                                Lambert / GGX / mirror / glass records with
 *                                emission instead of MDL arg blocks
 *   stk_build_bvh / stk_set_bvh  createBvhBuffer + BvhBuilder::build
 *                                (src/render/ptrender.cpp:460-532,
 *                                 src/render/bvh.cpp:237-254)
 *   stk_trace_closest[_device]   closestHit (shaders/raytracing.h:198-247)
 *   stk_trace_any[_device]       anyHit (shaders/raytracing.h:249-286)
 *   stk_render                   PtRender::drawFrame passes: path tracer,
 *                                reduction, accumulation, tonemap
 *                                (src/render/ptrender.cpp:641-913)
 *   stk_reset_accumulation       mPtIteration = 0 (src/render/ptrender.cpp:714-726)
 */
#ifndef STK_H
#define STK_H

#include <stddef.h>
#include <stdint.h>

#include "stk_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(_WIN32)
#define STK_API __declspec(dllexport)
#else
#define STK_API __attribute__((visibility("default")))
#endif

/* Result codes; the reference's Result{eOk,eFail,eOutOfMemory}
 * (include/render/common.h:53-58) widened with device/state/timeout cases. */
typedef enum stk_status {
    STK_OK = 0,
    STK_E_INVALID = 1, /* bad argument, size or layout */
    STK_E_OOM = 2,     /* device or host allocation failed */
    STK_E_DEVICE = 3,  /* a HIP call failed or no device is present */
    STK_E_STATE = 4,   /* call order violated (e.g. trace before geometry) */
    STK_E_TIMEOUT = 5  /* a bounded device wait expired */
} stk_status;

typedef struct stk_ctx stk_ctx;

/* ---- BVH build options (BvhBuilder::build, src/render/bvh.cpp:107-167) ---- */
typedef struct StkBvhOptions {
    uint32_t maxDepth;   /* root at depth 0; leaves at depth <= maxDepth (<= 32) */
    uint32_t numBins;    /* SAH bins per axis (0 = 32) */
    float traversalCost; /* SAH costs, reference uses 1 and 1 */
    float intersectionCost;
    /* > 0: spatial splits (the reference asks Embree for them,
     * src/render/bvh.cpp:80-105,154): where an object split's children
     * overlap by more than this fraction of the root box's area, a binned
     * spatial split may clip straddling triangles into both children (a
     * triangle then sits in several leaves; references grow by at most 30%).
     * 0 = object splits only.  Host builds of flat trees only. */
    float spatialSplitAlpha;
} StkBvhOptions;

/* Header of the packed BVH: root box and root reference. */
typedef struct StkBvhHeader {
    float lo[3];
    int32_t root; /* >= 0 packed node, < 0 leaf (~triangle), INT32_MIN empty */
    float hi[3];
    uint32_t numTriangles;
} StkBvhHeader;

/* trace flags */
/* Stackless skip-link walk in the reference's visit order, unpruned (closest)
 * or stopping at the first leaf in preorder (any).  Without the flag the
 * packed near-first (closest) or packed left-first (any) walks run; they
 * return the same hit records. */
#define STK_TRACE_REFERENCE_ORDER 0x1u
/* Closest hit over the packed binary BVH instead of the compressed 8-wide one
 * (same hit records; kept for measurement). */
#define STK_TRACE_BVH2 0x2u

/* ---- library / context ---------------------------------------------------- */
STK_API const char* stk_version(void);
/* "source-sha256 <hex>; offload-arch <arch>": a SHA-256 over the sources the
 * library was built from and their compile flags, so a caller can check that a
 * binary matches a source tree (strelka_amd/build.py: source_hash). */
STK_API const char* stk_build_info(void);
STK_API stk_status stk_create(int hipDevice, stk_ctx** out);
STK_API void stk_destroy(stk_ctx* ctx);
STK_API const char* stk_last_error(const stk_ctx* ctx);
STK_API stk_status stk_sync(stk_ctx* ctx);
/* The hipStream_t every call of this context is enqueued on. */
STK_API void* stk_stream(stk_ctx* ctx);

/* ---- scene input (deep-copied into device memory owned by the context) ---- */
STK_API stk_status stk_set_geometry(stk_ctx* ctx, const StkVertex* vb, uint32_t numVertices, const uint32_t* ib,
                                    uint32_t numIndices, const StkInstance* inst, uint32_t numInstances);
STK_API stk_status stk_set_lights(stk_ctx* ctx, const StkLight* lights, uint32_t numLights);
/* STK_E_INVALID for a record of unknown type or with a negative (or NaN)
 * index of refraction. */
STK_API stk_status stk_set_materials(stk_ctx* ctx, const StkMaterial* mats, uint32_t numMaterials);
/* Textures the materials' albedo, roughness, emission and normal texture
 * fields refer to (the reference's MDL materials sample theirs through
 * shaders/mdl_runtime.hlsl:173-207).  texels: numTexels RGBA32F values (4
 * floats each); every texture must lie inside them, be at most 16384 x 16384
 * and have a known wrap (STK_TEX_*).  Single-scene renders only (a batch
 * refuses textured materials). */
STK_API stk_status stk_set_textures(stk_ctx* ctx, const StkTexture* textures, uint32_t numTextures,
                                    const float* texels, uint64_t numTexels);
/* Builds the BVH on the host (binned SAH, one triangle per leaf, preorder plus
 * skip links) and uploads it.  opts may be NULL for defaults. */
STK_API stk_status stk_build_bvh(stk_ctx* ctx, const StkBvhOptions* opts);
/* The whole scene of a stk_scene container (stk_scene.h: generated, or built
 * with stk_scene_create_mesh / _instance / _light / _add_material): uploads its
 * geometry, instances, lights and materials and builds the flat BVH with opts
 * (NULL = defaults) - what oka::PtRender does with the oka::Scene handed to
 * setScene (include/render/ptrender.h:64, src/render/ptrender.cpp:656-667). */
#ifndef STK_SCENE_HANDLE
#define STK_SCENE_HANDLE
typedef struct stk_scene stk_scene;
#endif
STK_API stk_status stk_set_scene(stk_ctx* ctx, const stk_scene* scene, const StkBvhOptions* opts);
/* Builds the flat BVH on the device instead of the host (no reference
 * counterpart; the reference builds with Embree on the host,
 * src/render/bvh.cpp:107-167): a linear BVH (Morton-ordered centroids, binary
 * radix tree, one triangle per leaf) in the same skip-link preorder layout,
 * with the packed and compressed 8-wide layouts derived on the device.  Hits
 * are the reference walk's over that tree; its shape differs from
 * stk_build_bvh's SAH tree.  A tree deeper than the binary walk's stack (32)
 * is walked by the wide walk (closest hit, renders) or the stackless
 * skip-link walk (any hit); the binary-walk flag is then refused. */
STK_API stk_status stk_build_bvh_gpu(stk_ctx* ctx);
/* stk_build_bvh_gpu with a choice of topology: STK_GPU_BUILD_LBVH (the binary
 * radix tree above) or STK_GPU_BUILD_PLOC (parallel locally-ordered
 * clustering, Meister and Bittner 2018: clusters in Morton order merge with
 * the neighbour, within 32 positions, whose merged box has the least area;
 * a tree of lower area cost than the radix tree, built in the same time). */
#define STK_GPU_BUILD_LBVH 0u
#define STK_GPU_BUILD_PLOC 1u
STK_API stk_status stk_build_bvh_gpu_method(stk_ctx* ctx, uint32_t method);
/* Builds a two-level structure instead of the flat BVH (no reference
 * counterpart: the reference flattens every instance into one tree,
 * src/render/ptrender.cpp:460-532): one tree per distinct mesh (instance index
 * range) in object space and one over the instances' world boxes, for scenes
 * of many instances of few meshes.  Renders and closest / any-hit traces then
 * use it; hits are those of the reference walk (closestHit / anyHit) over the
 * structure expanded into one skip-link tree, the hit test running on the
 * world triangle getTriangle forms.  The reference-order and binary trace
 * flags and stk_get_bvh need the flat BVH (STK_E_STATE). */
STK_API stk_status stk_build_two_level(stk_ctx* ctx, const StkBvhOptions* opts);
/* No counterpart in the reference (it keeps only flattened world triangles,
 * src/render/ptrender.cpp:460-532).  enable != 0, after stk_build_two_level:
 * writes every instance's world triangles - the ones the two-level walks form
 * per triangle test from the object-space mesh and objectToWorld
 * (getTriangle, shaders/raytracing.h:134-153) - into device memory once, 48 B
 * per instanced triangle (18.8 M triangles: 905 MB), and the walks read them
 * instead; hits are bit-identical either way.  enable == 0 frees them.  A new
 * stk_build_two_level drops them.  STK_E_STATE without a two-level structure,
 * STK_E_INVALID beyond 2^31 instanced triangles. */
STK_API stk_status stk_two_level_world_triangles(stk_ctx* ctx, int enable);
/* Uses a caller-built skip-link BVH (e.g. the reference builder's output). */
STK_API stk_status stk_set_bvh(stk_ctx* ctx, const StkBvhNode* nodes, uint32_t numNodes);
/* Copies the skip-link BVH of the context out (numNodes receives the count;
 * nodes may be NULL to query the count). */
STK_API stk_status stk_get_bvh(stk_ctx* ctx, StkBvhNode* nodes, uint32_t capacity, uint32_t* numNodes);

/* What the context traces: the flat tree's header and depths, or a two-level
 * structure.  No counterpart in the reference (its BVH lives in a Vulkan
 * buffer, src/render/ptrender.cpp:516-532); for tests and tools that check
 * which walk a tree takes. */
typedef struct StkBvhInfo {
    StkBvhHeader header;
    uint32_t depth;       /* binary tree depth (root 0) */
    uint32_t wideDepth;   /* levels of the 8-wide tree (root = level 1) */
    uint32_t numNodes;    /* skip-link nodes */
    uint32_t numWide;     /* 8-wide nodes */
    uint32_t numEmitters; /* emitter triangles tested first by shadow rays */
    uint32_t twoLevel;    /* 1: instance tree over mesh trees */
    uint32_t gpuBuilt;    /* 1: built by stk_build_bvh_gpu* */
    uint32_t walkStack;   /* stack entries of the closest-hit kernel it takes (0: no wide walk) */
} StkBvhInfo;
STK_API stk_status stk_get_bvh_info(stk_ctx* ctx, StkBvhInfo* out);

/* ---- traversal ------------------------------------------------------------ */
/* Host buffers; synchronous. */
STK_API stk_status stk_trace_closest(stk_ctx* ctx, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags);
STK_API stk_status stk_trace_any(stk_ctx* ctx, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags);
/* stk_trace_closest without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  The rays and hits are staged through one
 * static device buffer of the library (52 MiB on each GPU it runs on, held
 * until the process ends) instead of buffers the context allocates, so at most
 * STK_NOALLOC_MAX_RAYS rays per call (more: STK_E_INVALID, nothing traced).
 * Same results as stk_trace_closest; returns when the hits are in `hits`.
 * Every context on one GPU shares that GPU's buffer: calls that overlap on one
 * GPU (from several threads or contexts) race. */
#define STK_NOALLOC_MAX_RAYS 1048576u
STK_API stk_status stk_trace_closest_noalloc(stk_ctx* ctx, const StkRay* rays, uint32_t n, StkHit* hits,
                                             uint32_t flags);
/* Device buffers; enqueued on the context stream, asynchronous. */
STK_API stk_status stk_trace_closest_device(stk_ctx* ctx, const StkRay* dRays, uint32_t n, StkHit* dHits,
                                            uint32_t flags);
/* No reference counterpart (the reference traces one camera ray per compute
 * thread, shaders/pathtracerMdl.hlsl:406-423): the same closest hits for an
 * image's rays, width x height in row-major order, traced with one wave per
 * 8x8 tile of pixels, so that a wave's rays stay close in the scene; hits are
 * written at the rays' own indices, identical to stk_trace_closest_device's.
 * flags as there: the wide, binary (STK_TRACE_BVH2) and two-level walks take
 * the tiles; the reference-order walk takes the rays in order. */
STK_API stk_status stk_trace_closest_image_device(stk_ctx* ctx, const StkRay* rays, uint32_t width, uint32_t height,
                                                  StkHit* hits, uint32_t flags);
STK_API stk_status stk_trace_any_device(stk_ctx* ctx, const StkRay* dRays, uint32_t n, StkHit* dHits,
                                        uint32_t flags);
/* Camera rays of generateCameraRay (shaders/pathtracerMdl.hlsl:165-202) for
 * every pixel, sample `sampleNum` of frame p->frameNumber, written in pixel
 * order into dRays (device, W*H rays).  Asynchronous. */
STK_API stk_status stk_generate_camera_rays_device(stk_ctx* ctx, const StkPathTracerParam* p, uint32_t sampleNum,
                                                   StkRay* dRays);

/* ---- rendering ------------------------------------------------------------ */
/* Progressive: traces p->pt.spp samples per pixel of frame p->pt.frameNumber,
 * reduces them, accumulates into the history (iteration counter kept by the
 * context, reset by stk_reset_accumulation or a dimension change) and
 * tonemaps.  outRgba receives W*H RGBA32F (tonemapped if enableTonemap,
 * otherwise the accumulated radiance); it may be a host or a device pointer
 * (outIsDevice).  Synchronous for host output, asynchronous otherwise. */
STK_API stk_status stk_render(stk_ctx* ctx, const StkRenderParams* p, float* outRgba, int outIsDevice);
/* stk_render into a caller-described image: a row pitch and RGBA32F or RGBA8
 * pixels (the Image handed to drawFrame, src/render/ptrender.cpp:893-906, and
 * Hydra render buffers, src/hdStrelka/RenderBuffer.cpp:78-81).  RGBA8 stores
 * round(clamp(v, 0, 1) * 255) per channel (NaN stores 0). */
#define STK_IMAGE_RGBA32F 0u
#define STK_IMAGE_RGBA8 1u
typedef struct StkImageDesc {
    void* data;          /* first byte of row 0 */
    uint32_t width, height;
    uint32_t pitchBytes; /* bytes from one row to the next; 0 = width * pixel size */
    uint32_t format;     /* STK_IMAGE_RGBA32F or STK_IMAGE_RGBA8 */
    int32_t onDevice;    /* data is a HIP device (or mapped host) pointer */
    uint32_t pad;
} StkImageDesc;
STK_API stk_status stk_render_image(stk_ctx* ctx, const StkRenderParams* p, const StkImageDesc* out);
STK_API stk_status stk_reset_accumulation(stk_ctx* ctx);
/* Watchdog (no reference counterpart: the reference has no failure detection).
 * With milliseconds > 0, a stk_render / stk_batch_render frame still running
 * that long after the call began is abandoned: a host thread sets an abort
 * word in pinned memory; the path tracer's one-wave bookkeeping kernels copy
 * it to the pass's device word at the start of every bounce, every later
 * kernel of the pass returns at entry, a persistent extend stops taking work
 * within 16 fetches per wave, and the accumulation step returns without
 * touching the image.  The call then returns STK_E_TIMEOUT with the
 * accumulated image and iteration count as they were before it (the output
 * buffer's contents are unspecified); a radiance cache keeps the optimiser
 * steps that finished before the abort and drops the rest.  While a deadline
 * is armed every render call waits for its frame.  0 (the default) disarms.
 * stk_trace_* calls are single bounded launches and are not watched. */
STK_API stk_status stk_set_watchdog(stk_ctx* ctx, uint32_t milliseconds);
/* Sets the abort word now, from any thread: the frame in flight (or, if none
 * is, the next one) is abandoned as above and its render call returns
 * STK_E_TIMEOUT; when that call had already returned (device output, no
 * deadline armed), the next render call undoes the frame's accumulation step
 * and returns STK_E_TIMEOUT without rendering.  With several pass streams
 * (which overlap consecutive calls) and device output, an abort can also cut
 * an earlier call's frame: the accumulation then restarts. */
STK_API stk_status stk_abort(stk_ctx* ctx);
/* Per-sample radiance of the last stk_render call, sampleBuffer layout
 * [(pixel*spp + s)*3 + c] (shaders/pathtracerMdl.hlsl:420-422). */
STK_API stk_status stk_get_samples(stk_ctx* ctx, float* out, size_t count);
/* The accumulated (not tonemapped) RGBA32F image. */
STK_API stk_status stk_get_accumulation(stk_ctx* ctx, float* out, size_t count);

/* Checkpoint / resume of the progressive state (no reference counterpart: the
 * reference keeps the accumulation image and mPtIteration only in GPU memory,
 * src/render/ptrender.cpp:714-726,839).  stk_checkpoint_save writes, to one
 * file: the accumulated image of the last stk_render (running mean, RGBA32F),
 * its iteration count and image layout; and, when the context has a radiance
 * cache, its configuration, step counter, FP32 weights, Adam moments and
 * moving averages.  userFrame is stored for the caller (the frameNumber to
 * continue from).  The file is written aside and renamed over path.
 * stk_checkpoint_load restores that state into a context whose scene (and, for
 * the cache section, whose cache: stk_cache_create with the same
 * configuration over the same scene bounds) the caller has set up again; the
 * next stk_render of the same image size continues the accumulation.  A
 * damaged or truncated file, or a cache of another configuration, is refused
 * (STK_E_INVALID / STK_E_STATE) before any state changes. */
typedef struct StkCheckpointInfo {
    uint32_t sections;  /* bit 0 accumulation, bit 1 cache */
    uint32_t iteration; /* frames accumulated */
    uint64_t pixels;    /* pixels of the accumulated image */
    uint64_t userFrame;
    uint32_t cacheStep; /* Adam steps taken */
    uint32_t pad;
} StkCheckpointInfo;
STK_API stk_status stk_checkpoint_save(stk_ctx* ctx, const char* path, uint64_t userFrame);
STK_API stk_status stk_checkpoint_load(stk_ctx* ctx, const char* path, StkCheckpointInfo* info);
/* Queue counters of the last wavefront pass (host copy of 8 words):
 * [0],[1] path queues of the last two bounces, [2] shadow rays of the last
 * bounce, [3] cache queries, [4] cache training records produced. */
STK_API stk_status stk_get_counters(stk_ctx* ctx, uint32_t out[8]);
/* Rays traced by the last render call: [0] closest-hit (extension) rays,
 * [1] shadow (any-hit) rays, over all its passes.  No reference counterpart
 * (a measurement aid). */
STK_API stk_status stk_get_ray_totals(stk_ctx* ctx, uint64_t out[2]);

/* ---- launch shape ----------------------------------------------------------
 * No counterpart in the reference (its compute passes use a fixed 16x16
 * workgroup, src/render/pathtracer.cpp).  blockSize 64/128/256 threads for the
 * wavefront kernels, raygen tile edge in pixels (1 = scanline order),
 * persistentBlocks > 0 makes the extend kernel a persistent grid of that many
 * blocks fetching rays from a work counter, useGraph captures one sample pass
 * as a hipGraph replayed for every pass (plain rendering only), and
 * pathsInFlight > 0 (at least 4096) traces a frame's paths in passes of that
 * many slots, so a pass's path state can stay in the last-level cache (plain
 * rendering only; the cache trains on whole frames), binaryWalk selects the
 * traversal.  Results do not depend on any of these. */
typedef struct StkLaunchConfig {
    uint32_t blockSize;
    uint32_t tile;
    uint32_t persistentBlocks;
    uint32_t useGraph;
    uint32_t pathsInFlight;
    uint32_t binaryWalk; /* 1: packed binary-BVH walks instead of the compressed 8-wide ones */
    /* threads per stripe of the extend / shade / connect grids (0: one per
     * possible queue entry); a capped grid loops over its stripe's entries,
     * so late bounces with short queues launch few blocks */
    uint32_t stripeThreads;
    /* 1..4 (0 means 1): with n > 1, pass k of a frame runs on stream k % n,
     * each stream with its own path state, so that one pass's last bounces
     * overlap the next passes' first ones; the passes still add into the
     * image in sample order; with the radiance cache a pass's inference and
     * training also wait for the previous pass's (no graph replay then) */
    uint32_t passStreams;
    /* n > 0: plain rendering runs each frame as one persistent launch of n
     * workgroups of 256 threads over every view and sample: a workgroup takes
     * chunks of frameChunk slots from a counter and runs all of a chunk's
     * samples (raygen, bounces, shadow rays, reduction) with queues of its own,
     * stages separated by workgroup barriers (no grid-wide wait); the context
     * stream only (passStreams and useGraph do not apply).  0: one launch per
     * stage and bounce.  Renders with the radiance cache always use the
     * per-stage launches.  stk_get_counters is not kept in this mode. */
    uint32_t frameBlocks;
    /* slots per chunk of the frame launch: a multiple of 256 up to 8192 (0:
     * 256); a chunk larger than the workgroup keeps its lanes busy in the
     * late bounces */
    uint32_t frameChunk;
} StkLaunchConfig;
STK_API stk_status stk_set_launch_config(stk_ctx* ctx, const StkLaunchConfig* lc);
STK_API stk_status stk_get_launch_config(stk_ctx* ctx, StkLaunchConfig* lc);
/* Properties of the context's device (for sizing launches). */
typedef struct StkDeviceInfo {
    char name[64];
    char arch[64]; /* e.g. "gfx950:sramecc+:xnack-" */
    uint32_t computeUnits;
    uint32_t wavefrontSize;
    uint64_t globalMemBytes;
    uint32_t l2Bytes;
    uint32_t clockKhz;
    uint32_t memClockKhz;
    uint32_t memBusWidth;
} StkDeviceInfo;
STK_API stk_status stk_device_info(stk_ctx* ctx, StkDeviceInfo* out);

/* ---- batched scenes --------------------------------------------------------
 * Many independent scenes, each with its own view, rendered together: every
 * wavefront kernel of a sample pass runs once over the path slots of all
 * views.  The reference renders one scene per PtRender; a host loop over
 * PtRender::drawFrame (src/render/ptrender.cpp:641-913) per scene is what a
 * batch replaces.  Buffers are copied by stk_batch_add_scene; instance
 * indexOffset/materialId/lightId are relative to the scene's own arrays. */
typedef struct StkBatchScene {
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
    StkPathTracerParam camera; /* matrices and dimension of the view */
    float sky[4];              /* constant sky radiance */
} StkBatchScene;
typedef struct StkBatchViewInfo {
    uint32_t pixelBase; /* first pixel of the view in the batched image */
    uint32_t width, height;
    uint32_t numTriangles;
    uint32_t bvhDepth;
    uint32_t numScenes;
    uint32_t pad[2];
} StkBatchViewInfo;
STK_API stk_status stk_batch_clear(stk_ctx* ctx);
STK_API stk_status stk_batch_add_scene(stk_ctx* ctx, const StkBatchScene* scene, uint32_t* index);
/* Builds every scene's BVH (host threads) and uploads the merged arrays. */
STK_API stk_status stk_batch_commit(stk_ctx* ctx, const StkBvhOptions* opts);
STK_API stk_status stk_batch_info(stk_ctx* ctx, uint32_t scene, StkBatchViewInfo* out);
/* stk_render over every committed view: p supplies spp, maxDepth,
 * frameNumber, sampling, rrStartDepth, accumulation and tonemap settings
 * (its camera, dimension and sky are ignored; cacheMode must be 0).  outRgba
 * receives the views one after another (view i at pixelBase_i, width_i x
 * height_i RGBA32F). */
STK_API stk_status stk_batch_render(stk_ctx* ctx, const StkRenderParams* p, float* outRgba, int outIsDevice);

/* ---- radiance cache -------------------------------------------------------
 * No counterpart in the reference renderer: an online-trained cache of
 * outgoing radiance (multi-resolution hash grid of 16 levels x 2 FP16
 * features feeding a 64-wide MLP with 4 ReLU hidden layers on MFMA, FP16
 * weights/activations, FP32 accumulation, Adam).  The plain path tracer
 * (cacheMode 0) stays available as the fallback and as the reference the
 * cache's bias is measured against. */
typedef struct StkCacheConfig {
    uint32_t log2TableSize;  /* hash-table entries per level: 2^19 */
    uint32_t baseResolution; /* coarsest grid resolution: 16 */
    uint32_t maxResolution;  /* finest grid resolution: 2048 */
    uint32_t seed;           /* weight initialisation */
    float learningRate;      /* Adam: 1e-2 */
    float beta1, beta2, epsilon;
    uint32_t maxBatch;       /* capacity of one training step, in records */
    float emaDecay;          /* queries read an exponential moving average of the
                                weights with this decay (0 = the raw weights) */
    /* learning-rate decay: step k (from 1) uses
     * max(minLearningRate, learningRate * 2^(-(k - 1) / lrHalfLife)); 0 = none */
    float lrHalfLife;
    float minLearningRate;
} StkCacheConfig;

typedef struct StkCacheInfo {
    float lo[3];
    float scale;              /* normalised position = (p - lo) * scale */
    uint32_t resolution[16];
    uint32_t dense[16];       /* 1 if the level is indexed densely */
    uint32_t log2TableSize;
    uint32_t gridParams;      /* 16 * 2^log2TableSize * 2 */
    uint32_t mlpParams;       /* 4 * 64 * 64 + 64 * 16 */
    uint32_t step;            /* Adam steps taken */
} StkCacheInfo;

/* Creates (or re-creates) the cache of a context; the scene bounds come from
 * the BVH, so a BVH must be set.  cfg may be NULL for the defaults above. */
STK_API stk_status stk_cache_create(stk_ctx* ctx, const StkCacheConfig* cfg);
STK_API stk_status stk_cache_destroy(stk_ctx* ctx);
STK_API stk_status stk_cache_info(stk_ctx* ctx, StkCacheInfo* out);
/* Predicted radiance (float4 rgb + 0 per record), device buffers, async. */
STK_API stk_status stk_cache_infer(stk_ctx* ctx, const StkCacheRecord* dRecs, uint32_t n, float* dOut);
/* One training step (backward + Adam) on n device records; lossOut (host,
 * may be NULL) receives the mean relative-L2 loss.  Synchronous. */
STK_API stk_status stk_cache_train_step(stk_ctx* ctx, const StkCacheRecord* dRecs, uint32_t n, float* lossOut);
/* Gradients of the summed relative-L2 loss for n device records, copied to
 * host arrays (grid: gridParams floats, mlp: mlpParams floats, W[l][in][out]);
 * parameters are left unchanged.  Synchronous. */
STK_API stk_status stk_cache_gradients(stk_ctx* ctx, const StkCacheRecord* dRecs, uint32_t n, float* gridGrad,
                                       float* mlpGrad, float* lossSum);
/* FP32 master parameters in/out (host arrays of gridParams / mlpParams). */
STK_API stk_status stk_cache_get_params(stk_ctx* ctx, float* grid, float* mlp);
STK_API stk_status stk_cache_set_params(stk_ctx* ctx, const float* grid, const float* mlp);
/* The records the last cached render produced (no reference counterpart):
 * which = 0 the cache queries of its last pass (targets unused; normal.w holds
 * the path slot), 1 the training records of the frame with their targets.
 * out may be NULL to query *n; waits for the render. */
STK_API stk_status stk_cache_last_records(stk_ctx* ctx, uint32_t which, StkCacheRecord* out, uint32_t capacity,
                                          uint32_t* n);

/* ---- device memory and timing helpers for callers without a HIP binding --- */
STK_API stk_status stk_device_alloc(stk_ctx* ctx, size_t bytes, void** out);
/* Pinned host memory the device can read and write directly (the host-visible
 * buffers of ResourceManager::createBuffer + getMappedMemory,
 * include/resourcemanager/resourcemanager.h:31-66). */
STK_API stk_status stk_host_alloc(stk_ctx* ctx, size_t bytes, void** out);
STK_API stk_status stk_host_free(stk_ctx* ctx, void* ptr);
STK_API stk_status stk_device_free(stk_ctx* ctx, void* ptr);
STK_API stk_status stk_copy_to_device(stk_ctx* ctx, void* dst, const void* src, size_t bytes);
STK_API stk_status stk_copy_to_host(stk_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy through a plain streaming kernel (bytes a multiple of
 * 16); its known traffic calibrates the memory counters of a profiler. */
STK_API stk_status stk_device_copy(stk_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Records a HIP event on the context stream (slot 0..15). */
STK_API stk_status stk_event_record(stk_ctx* ctx, int slot);
/* Milliseconds between two recorded slots (waits for the second). */
STK_API stk_status stk_event_elapsed(stk_ctx* ctx, int slotBegin, int slotEnd, float* ms);

/* ---- files (host only) ----------------------------------------------------
 * No counterpart in the reference (its BVH lives in memory, frames go out as
 * PNG through HioImage, src/hdRunner/main.cpp:339-356).  BVH file: a 16-byte
 * header {"STKBVH01", node count, 0} and the skip-link StkBvhNode records that
 * stk_set_bvh accepts.  PFM: RGB float32, little-endian, bottom row first.
 * OpenEXR: uncompressed float scanlines. */
STK_API stk_status stk_bvh_save_nodes(const StkBvhNode* nodes, uint32_t numNodes, const char* path);
/* nodes may be NULL to query numNodes; the file must hold exactly the header
 * and numNodes records. */
STK_API stk_status stk_bvh_load_nodes(const char* path, StkBvhNode* nodes, uint32_t capacity, uint32_t* numNodes);
/* rgba: width x height RGBA32F, top row first (as stk_render writes it). */
STK_API stk_status stk_image_write_pfm(const char* path, const float* rgba, uint32_t width, uint32_t height);
/* OpenEXR 2.0 scanline file, uncompressed, channels B, G, R as 32-bit floats
 * (alpha dropped), top row first: the lossless format image tools read for
 * converged references.  rgba as for stk_image_write_pfm. */
STK_API stk_status stk_image_write_exr(const char* path, const float* rgba, uint32_t width, uint32_t height);
/* Reads an RGB PFM into RGBA32F rows (top row first, alpha 1); rgba may be
 * NULL to query the size. */
STK_API stk_status stk_image_read_pfm(const char* path, float* rgba, size_t capacityFloats, uint32_t* width,
                                      uint32_t* height);

/* ---- host-only BVH tools (no device needed) ------------------------------ */
typedef struct stk_bvh stk_bvh;
/* Flattens instances to world-space triangles (createBvhBuffer,
 * src/render/ptrender.cpp:460-512) and builds the BVH. */
STK_API stk_status stk_bvh_build_host(const StkVertex* vb, uint32_t numVertices, const uint32_t* ib,
                                      uint32_t numIndices, const StkInstance* inst, uint32_t numInstances,
                                      const StkBvhOptions* opts, stk_bvh** out);
/* Derives the packed form from a skip-link node array. */
STK_API stk_status stk_bvh_from_nodes(const StkVertex* vb, uint32_t numVertices, const uint32_t* ib,
                                      uint32_t numIndices, const StkInstance* inst, uint32_t numInstances,
                                      const StkBvhNode* nodes, uint32_t numNodes, stk_bvh** out);
STK_API void stk_bvh_free(stk_bvh* b);
STK_API uint32_t stk_bvh_num_nodes(const stk_bvh* b);
STK_API const StkBvhNode* stk_bvh_nodes(const stk_bvh* b);
STK_API uint32_t stk_bvh_num_packed(const stk_bvh* b);
STK_API const StkPackedNode* stk_bvh_packed(const stk_bvh* b);
STK_API uint32_t stk_bvh_num_triangles(const stk_bvh* b);
STK_API const StkTriangle* stk_bvh_triangles(const stk_bvh* b);
STK_API StkBvhHeader stk_bvh_header(const stk_bvh* b);
STK_API uint32_t stk_bvh_depth(const stk_bvh* b);
/* The compressed 8-wide layout the GPU traversal walks (StkWideNode), its
 * triangles in slot order (stk_bvh_num_triangles of them), its number of
 * inner-node levels, and the number of emitter triangles. */
STK_API uint32_t stk_bvh_num_wide(const stk_bvh* b);
STK_API const StkWideNode* stk_bvh_wide_nodes(const stk_bvh* b);
STK_API const StkTriangle* stk_bvh_wide_triangles(const stk_bvh* b);
STK_API uint32_t stk_bvh_wide_depth(const stk_bvh* b);
STK_API uint32_t stk_bvh_num_emitters(const stk_bvh* b);

/* ---- two-level structure on the host (no device needed) ------------------------
 * The same build stk_build_two_level runs, for tools and tests: a binary and a
 * compressed 8-wide tree per distinct mesh (instance index range) in object
 * space, and over the instances' world boxes.  No reference counterpart: the
 * reference flattens instances (src/render/ptrender.cpp:460-532). */
typedef struct stk_tlas stk_tlas;
STK_API stk_status stk_tlas_build_host(const StkVertex* vb, uint32_t numVertices, const uint32_t* ib,
                                       uint32_t numIndices, const StkInstance* inst, uint32_t numInstances,
                                       const StkBvhOptions* opts, stk_tlas** out);
STK_API void stk_tlas_free(stk_tlas* t);
STK_API uint32_t stk_tlas_num_meshes(const stk_tlas* t);
/* mesh of instance i, STK_INVALID_INDEX for an instance without triangles */
STK_API uint32_t stk_tlas_instance_mesh(const stk_tlas* t, uint32_t i);
/* a mesh's binary skip-link tree (object-space boxes; leaves: primId in
 * minBounds[0], instId 0) */
STK_API const StkBvhNode* stk_tlas_mesh_nodes(const stk_tlas* t, uint32_t mesh, uint32_t* count);
/* the binary skip-link tree over the instances (leaves: instId = instance) */
STK_API const StkBvhNode* stk_tlas_top_nodes(const stk_tlas* t, uint32_t* count);
STK_API const StkWideNode* stk_tlas_top_wide(const stk_tlas* t, uint32_t* count);
STK_API const StkTlasLeaf* stk_tlas_top_leaves(const stk_tlas* t, uint32_t* count);
/* every mesh's wide tree in one array, and their leaf triangles */
STK_API const StkWideNode* stk_tlas_mesh_wide(const stk_tlas* t, uint32_t* count);
STK_API const StkBlasTriangle* stk_tlas_mesh_triangles(const stk_tlas* t, uint32_t* count);
/* per instance: its leaf's preorder index in the top tree; 16 floats: the
 * rows of the inverse transform the walk uses, then {pad, 0, 0, 0} */
STK_API const uint32_t* stk_tlas_instance_keys(const stk_tlas* t);
STK_API const float* stk_tlas_instance_inverses(const stk_tlas* t);
/* world triangles of emitter instances (leaf = the mesh-tree leaf) */
STK_API const StkTriangle* stk_tlas_emitters(const stk_tlas* t, uint32_t* count);
STK_API void stk_tlas_depths(const stk_tlas* t, uint32_t* top, uint32_t* mesh);

#ifdef __cplusplus
}
#endif
#endif
