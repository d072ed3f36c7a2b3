// kernels.hpp - host-side launchers of the HIP kernels (kernels.hip,
// pathtrace.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/stk/stk.h"
#include "cache.hpp"
#include "trace_dev.hpp"

namespace stk {

// One scene of a (possibly batched) render: the scenes' BVHs, triangles,
// instances, lights and materials are concatenated into one forest; this
// record locates a scene's part and carries its view.
struct SceneDesc {
    StkBvhHeader hdr;        // root rebased into the merged packed-node array
    uint32_t refRoot;        // first node of the scene in the merged skip-link array
    uint32_t wideRoot;       // root in the merged wide-node array (STK_INVALID_INDEX: use the packed walk)
    uint32_t emitBase, emitCount; // the scene's emitter triangles in the merged emitter array
    uint32_t lightBase, lightCount;
    uint32_t pixelBase;      // first pixel of the view in the merged image
    uint32_t slotBase;       // first path slot (views are padded to whole tiles)
    uint32_t width, height;
    uint32_t tilesX;         // tiles per row for this view's tile edge
    float sky[4];
    StkPathTracerParam cam;  // camera of the view (dimension = width, height)
};

struct SceneView {
    BvhView bvh; // merged forest (bvh.hdr unused)
    const StkVertex* vb;
    const uint32_t* ib;
    const StkInstance* inst;
    const StkLight* lights;
    const StkMaterial* mats;
    uint32_t numMats;
    const StkTexture* textures; // albedo textures (single-scene renders)
    uint32_t numTextures;
    const float4* texels;
    const SceneDesc* scenes;
    uint32_t numScenes;
    uint32_t totalSlots;
    uint32_t tile; // raygen tile edge in pixels (1 = scanline order)
};

// Per-pass values read from device memory, so that a captured graph of one
// pass can be replayed for every sample and frame.
struct PassState {
    uint32_t frameNumber;
    uint32_t sampleNum;
    uint32_t trainOffset;
    uint32_t numTrainPaths;
    uint32_t slotBegin; // the pass covers global slots [slotBegin, slotBegin + slotCount)
    uint32_t slotCount;
};

// Ray queues are kStripes segments of segCap entries each (see pathtrace.hip).
constexpr uint32_t kStripes = 64;
constexpr int kQueueKinds = 3; // path queue 0, path queue 1, shadow queue

// Entries per stripe segment for a pass of n slots produced by 256-thread
// blocks (block b appends to segment b % kStripes, at most 256 entries).
inline uint32_t stripeSegment(size_t n) { return (uint32_t)(((n + 255) / 256 + kStripes - 1) / kStripes * 256); }

// Device arrays of one wavefront path-tracing pass (one sample of the pixels of
// the pass's slots).  Per-slot arrays are indexed by the slot's position in
// the pass; hit, hitPrim and the queues by queue entry (stripe * segCap + i).
struct PathState {
    uint32_t capacity = 0;      // slots
    size_t queueCapacity = 0;   // entries of each queue-indexed array
    uint32_t segCap = 0;        // entries per stripe segment in the current pass
    uint32_t* stripes = nullptr; // [kQueueKinds][kStripes] entries per segment
    uint32_t* work = nullptr;    // [kStripes] persistent-extend fetch counters
    float4* rayO = nullptr;  // origin.xyz, tmax
    float4* rayD = nullptr;  // direction.xyz
    float4* thr = nullptr;   // throughput.rgb
    float4* col = nullptr;   // radiance.rgb
    uint32_t* rng = nullptr; // xorshift state
    uint32_t* scene = nullptr; // scene of the slot
    uint32_t* pixel = nullptr; // global pixel of the slot (0xFFFFFFFF: padding)
    float4* hit = nullptr;   // u, v, t, instId bits (queue order)
    uint32_t* hitPrim = nullptr;
    uint32_t* queue[2] = {nullptr, nullptr};
    uint32_t* shadowQ = nullptr;
    float4* shO = nullptr;
    float4* shD = nullptr;
    float4* shCd = nullptr; // NEE contribution, diffuse part
    float4* shCg = nullptr; // NEE contribution, glossy part
    uint32_t* counters = nullptr; // [0],[1] queue totals, [2] shadow total, [3] queries, [4] train records,
                                  // [5] abort word of this pass (1: the pass's kernels return at entry),
                                  // [8..11] two 64-bit ray totals of the render call (closest, shadow)
    PassState* pass = nullptr;
    // Watchdog: a host word in pinned memory (nonzero: abandon the frame),
    // read by the one-wave bookkeeping kernels into counters[5] and, when
    // set, into abortAny (a device word of the context read by the
    // accumulation step after every pass stream has joined).
    const uint32_t* abortHost = nullptr;
    uint32_t* abortAny = nullptr;
};

// Index of the pass's abort word in PathState::counters.
constexpr int kAbortCounter = 5;

// Radiance-cache hooks of one pass.  Training paths are the slots with
// slot % trainStride == trainOffset: they run to maxDepth and record every
// surface vertex (6 float4: position+roughness, normal, outgoing direction,
// radiance before the vertex, throughput at the vertex, radiance when the path
// left the vertex).  Other paths query the cache at the first vertex of depth
// >= queryDepth whose roughness is >= queryMinRough: they sample its direct
// light as usual and end there, the cache supplying the rest.
struct PtCache {
    uint32_t mode = 0; // 0 off, 1 query + train, 2 query only, 3 train only, 4 warm-up
    uint32_t queryDepth = 2;
    float queryMinRough = 0.0f;
    uint32_t trainSteps = 1; // optimiser steps per frame over the frame's records
    uint32_t trainSuffix = 0; // > 0: training paths end with a query at depth >= queryDepth + trainSuffix
    uint32_t queryDirect = 0; // 1: a query covers its vertex's direct light (no shadow ray there)
    uint32_t trainStride = 1;
    uint32_t trainCap = 0;
    uint32_t maxTrainPaths = 0;
    CacheRecord* queryRecs = nullptr; // [slots]
    float4* queryOut = nullptr;       // [slots]
    float4* trainVerts = nullptr;     // [maxTrainPaths * maxDepth * 6]
    uint32_t* trainCount = nullptr;   // [maxTrainPaths]
    CacheRecord* trainBatch = nullptr; // [trainCap]
};

struct PtLaunch {
    uint32_t maxDepth;
    uint32_t spp;
    uint32_t samplingType; // stratifiedSamplingType
    uint32_t rrStartDepth;
    uint32_t refCompat;
    float invSpp;
    uint32_t passSlots; // slots per pass (grid size of the per-slot kernels), <= PathState capacity
    PtCache cache;
};

// Launch shape of the wavefront kernels (chosen by the autotuner).
struct LaunchConfig {
    uint32_t blockSize = 256;      // 64, 128 or 256 threads
    uint32_t tile = 8;             // raygen tile edge (1 = scanline)
    uint32_t persistentBlocks = 0; // extend grid for persistent waves (0: one lane per ray)
    uint32_t pathsInFlight = 0;    // slots per pass (0: every slot of the frame in one pass)
    uint32_t binaryWalk = 0;       // 1: packed BVH2 walks instead of the compressed 8-wide ones
    uint32_t stripeThreads = 0;    // threads per stripe of the consumer grids (0: one per entry)
    uint32_t passStreams = 1;      // n > 1: pass k runs on stream k % n
    uint32_t frameBlocks = 0;      // n > 0: one persistent launch of n workgroups per frame (plain rendering)
    uint32_t frameChunk = 0;       // slots per chunk of that launch (a multiple of 256; 0: 256)
};

// Stack entries per lane of the wide closest-hit kernel a flat tree takes (11
// for trees of at most 12 wide levels, else 16); 0 when the tree has no wide
// walk (two-level structures, or no wide tree).
uint32_t closestWalkStack(const BvhView& bvh);
// The two-level structure's world triangles (BvhView::worldTris): instance i's
// count[i] mesh leaves from blasTris[meshBase[i]] on, through its
// objectToWorld, to worldTris[worldBase[i]] on.
hipError_t launchWorldTriangles(const StkInstance* inst, const float4* blasTris, const uint32_t* meshBase,
                                const uint32_t* count, const uint64_t* worldBase, uint32_t numInstances,
                                float4* worldTris, hipStream_t s);

// flags: STK_TRACE_REFERENCE_ORDER (skip-link walk), STK_TRACE_BVH2 (packed
// binary walk instead of the wide one; closest hit only).
hipError_t launchTraceClosest(const BvhView& bvh, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags,
                              hipStream_t s, uint32_t width = 0);
hipError_t launchTraceAny(const BvhView& bvh, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags,
                          hipStream_t s);
hipError_t launchCameraRays(const StkPathTracerParam& p, uint32_t sampleNum, StkRay* rays, hipStream_t s);

// Writes the pass state (a kernel, so it can precede a graph replay).
hipError_t launchSetPass(PathState& ps, const PassState& v, hipStream_t s);
// Clears the 64-bit ray totals (before the first pass of a render call).
hipError_t launchClearRayTotals(PathState& ps, hipStream_t s);
// One sample pass of the wavefront path tracer over every view: raygen, the
// bounce loop and the finalize step, which adds invSpp * radiance into
// accum[pixel] (and stores per-sample radiance when samples is set).  With
// lp.cache.mode != 0, cb must be the context's cache: queries are answered
// after the bounce loop and one training step runs on the pass's records.
// finalizeAfter (optional): the finalize step waits for this event (the
// previous pass's finalize on another stream); finalizeDone (optional) is
// recorded after it.  cacheAfter (optional): the cache's inference and
// training steps wait for it (the previous pass's training on another
// stream); cacheDone (optional) is recorded after them.
hipError_t launchPathTracePass(const SceneView& sc, const PtLaunch& lp, const LaunchConfig& lc, PathState& ps,
                               float4* accum, float* samples, CacheBuffers* cb, hipStream_t s,
                               hipEvent_t finalizeAfter = nullptr, hipEvent_t finalizeDone = nullptr,
                               hipEvent_t cacheAfter = nullptr, hipEvent_t cacheDone = nullptr);
// A whole frame (every view, lp.spp samples) in one persistent launch of
// lc.frameBlocks workgroups (after a one-wave reset of the chunk counter):
// each workgroup takes chunks of lc.frameChunk slots and runs all their
// samples with queues private to it, adding into accum in sample order.  ps
// must hold frameBlocks x frameChunk slots and queue entries; plain rendering only (cache
// mode 0).  The image equals the per-stage passes' bit for bit.
hipError_t launchPathTraceFrame(const SceneView& sc, const PtLaunch& lp, const LaunchConfig& lc, PathState& ps,
                                float4* accum, float* samples, uint32_t frameNumber, uint32_t totalSlots,
                                hipStream_t s);
// Plain 16-B-per-lane copy kernel (bytes must be a multiple of 16).
hipError_t launchCopy(void* dst, const void* src, size_t bytes, hipStream_t s);
// RGBA32F pixels to a tightly packed RGBA32F (format 0) or RGBA8 (format 1)
// image; RGBA8 stores round(clamp(v, 0, 1) * 255).
hipError_t launchPackImage(const float4* src, uint32_t npix, uint32_t format, void* dst, hipStream_t s);
// Zeroes n float4 (a kernel, so the call can be captured in a graph).
hipError_t launchZero(float4* p, size_t n, hipStream_t s);
// Running-mean accumulation (shaders/accumulation.hlsl:86-97) + tonemap
// (shaders/tonemap.hlsl:66-88) in one pass.
// abortAny (may be null): when it holds a nonzero word the step returns
// without touching the history or the output (a frame the watchdog abandoned).
hipError_t launchAccumulateTonemap(const float4* cur, float4* history, float4* out, uint32_t npix,
                                   uint32_t iteration, int accumulate, int tonemap, uint32_t tonemapType,
                                   const uint32_t* abortAny, hipStream_t s);

} // namespace stk
