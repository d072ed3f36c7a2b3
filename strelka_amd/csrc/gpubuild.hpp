// gpubuild.hpp - BVH construction on the GPU (stk_build_bvh_gpu).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>

#include "../../include/stk/stk.h"

namespace stk {

// Buffers the builder asks its caller for: the outputs the walks read (kept)
// and temporaries (released after the build).
enum GpuBuildSlot {
    kGbRefNodes = 0, // StkBvhNode, device form (leaves: triangle index in maxBounds[0])
    kGbPacked,       // StkPackedNode
    kGbTris,         // StkTriangle, packed-walk order
    kGbWide,         // StkWideNode
    kGbWideTris,     // StkTriangle, wide slot order
    kGbEmitters,     // StkTriangle of emitter instances
    kGbFirstTemp     // temporaries from here on
};

struct GpuBvhResult {
    uint32_t numNodes = 0, numPacked = 0, numTris = 0, numWide = 0, numEmit = 0;
    uint32_t depth = 0;     // levels of the binary tree (root = 0)
    uint32_t wideDepth = 0; // levels of the wide tree
    uint32_t rounds = 0;    // clustering rounds (kGpuBuildPloc)
    StkBvhHeader header{};
};

// Topology of the device-built tree.
enum GpuBuildMethod : uint32_t {
    kGpuBuildLbvh = STK_GPU_BUILD_LBVH, // binary radix tree over the Morton order
    kGpuBuildPloc = STK_GPU_BUILD_PLOC, // locally-ordered clustering over the Morton order
};

// Builds a BVH over every instance's world triangles (63-bit Morton codes of
// the centroids, radix sort, then the binary radix tree from longest common
// prefixes with bottom-up boxes (Karras 2012) or locally-ordered clustering
// (Meister and Bittner 2018), by `method`) and emits the skip-link preorder,
// packed and compressed 8-wide layouts of the host builder's format.
// instStart[i] is the first triangle of instance i (host array of ninst + 1
// entries); the device arrays are the context's geometry.
hipError_t gpuBuildBvh(const StkVertex* vb, const uint32_t* ib, const StkInstance* inst, const uint32_t* instStart,
                       uint32_t ninst, uint32_t ntri, uint32_t method,
                       const std::function<void*(int, size_t)>& alloc, GpuBvhResult& out, std::string& err,
                       hipStream_t s);

} // namespace stk
