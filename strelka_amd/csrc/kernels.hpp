// kernels.hpp - host-side launchers of the HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/stk/stk.h"
#include "cache.hpp"
#include "trace_dev.hpp"

namespace stk {

// Device arrays of one wavefront path-tracing pass (one sample of every pixel).
struct PathState {
    uint32_t capacity = 0;
    float4* rayO = nullptr;  // origin.xyz, tmax
    float4* rayD = nullptr;  // direction.xyz
    float4* thr = nullptr;   // throughput.rgb
    float4* col = nullptr;   // radiance.rgb
    uint32_t* rng = nullptr; // xorshift state
    float4* hit = nullptr;   // u, v, t, instId bits (queue order)
    uint32_t* hitPrim = nullptr;
    uint32_t* queue[2] = {nullptr, nullptr};
    uint32_t* shadowQ = nullptr;
    float4* shO = nullptr;
    float4* shD = nullptr;
    float4* shCd = nullptr; // NEE contribution, diffuse part
    float4* shCg = nullptr; // NEE contribution, glossy part
    uint32_t* counters = nullptr; // [0],[1] queue counts, [2] shadow count, [3..7] work counters
};

struct SceneView {
    BvhView bvh;
    const StkVertex* vb;
    const uint32_t* ib;
    const StkInstance* inst;
    const StkLight* lights;
    uint32_t numLights;
    const StkMaterial* mats;
    uint32_t numMats;
};

// Radiance-cache hooks of one pass.  Training paths are the slots with
// slot % trainStride == trainOffset: they run to maxDepth and record every
// surface vertex (5 float4: position+roughness, normal, outgoing direction,
// radiance before the vertex, throughput at the vertex).  Other paths query
// the cache at vertex depth queryDepth and end there.
struct PtCache {
    uint32_t mode = 0; // 0 off, 1 query + train, 2 query only, 3 train only
    uint32_t queryDepth = 2;
    uint32_t trainStride = 1;
    uint32_t trainOffset = 0;
    uint32_t trainCap = 0;
    uint32_t numTrainPaths = 0;
    CacheRecord* queryRecs = nullptr; // [npix]
    float4* queryOut = nullptr;       // [npix]
    float4* trainVerts = nullptr;     // [numTrainPaths * maxDepth * 5]
    uint32_t* trainCount = nullptr;   // [numTrainPaths]
    CacheRecord* trainBatch = nullptr; // [trainCap]
};

struct PtLaunch {
    StkPathTracerParam pt;
    float sky[4];
    uint32_t rrStartDepth;
    uint32_t refCompat;
    uint32_t sampleNum; // 0 .. spp-1
    float invSpp;
    PtCache cache;
};

hipError_t launchTraceClosest(const BvhView& bvh, const StkRay* rays, uint32_t n, StkHit* hits, bool refOrder,
                              hipStream_t s);
hipError_t launchTraceAny(const BvhView& bvh, const StkRay* rays, uint32_t n, StkHit* hits, bool refOrder,
                          hipStream_t s);
hipError_t launchCameraRays(const StkPathTracerParam& p, uint32_t sampleNum, StkRay* rays, hipStream_t s);

// One sample pass of the wavefront path tracer; accumulates invSpp * radiance
// into accum (float4 per pixel) and optionally stores per-sample radiance.
// With lp.cache.mode != 0, cb must be the context's cache: queries are answered
// after the bounce loop and one training step runs on the pass's records.
hipError_t launchPathTracePass(const SceneView& sc, const PtLaunch& lp, PathState& ps, float4* accum,
                               float* samples, CacheBuffers* cb, hipStream_t s);
// Plain 16-B-per-lane copy kernel (bytes must be a multiple of 16).
hipError_t launchCopy(void* dst, const void* src, size_t bytes, hipStream_t s);
// Zeroes n float4 (a kernel, so the call can be captured in a graph).
hipError_t launchZero(float4* p, size_t n, hipStream_t s);
// Running-mean accumulation (shaders/accumulation.hlsl:86-97) + tonemap
// (shaders/tonemap.hlsl:66-88) in one pass.
hipError_t launchAccumulateTonemap(const float4* cur, float4* history, float4* out, uint32_t npix,
                                   uint32_t iteration, int accumulate, int tonemap, uint32_t tonemapType,
                                   hipStream_t s);

} // namespace stk
