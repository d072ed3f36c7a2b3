// bvh.hpp - host BVH builder and packer.
#pragma once

#include <string>
#include <vector>

#include "../../include/stk/stk.h"
#include "geom.hpp"

namespace stk {

// One world-space triangle of the BVH input; the three positions of
// BVHInputPosition (include/render/bvh.h:31-36) with their (instId, primId).
struct BvhInputTri {
    F3 v[3];
    uint32_t instId;
    uint32_t primId;
};

struct Bvh {
    std::vector<StkBvhNode> nodes;     // reference skip-link layout, preorder
    std::vector<StkPackedNode> packed; // BVH2 nodes of the GPU traversal
    std::vector<StkTriangle> tris;     // pre-transformed triangles, leaf order
    std::vector<uint32_t> leafToTri;   // preorder node index -> triangle index
    StkBvhHeader header{};
    uint32_t depth = 0;
    // compressed 8-wide layout (buildWide)
    std::vector<StkWideNode> wide;     // node 0 is the root
    std::vector<StkTriangle> wideTris; // leaf triangles in wide-node slot order
    std::vector<StkTriangle> emitters; // triangles of emitter instances (lightId != -1)
    uint32_t wideDepth = 0;            // inner-node levels of the wide tree
};

// createBvhBuffer (src/render/ptrender.cpp:460-512): world-space positions of
// every instance's triangles, primId = triangle index within its instance.
bool flattenInstances(const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni, const StkInstance* inst,
                      uint32_t ninst, std::vector<BvhInputTri>& out, std::string& err);

// Binned-SAH build, one triangle per leaf, preorder + skip links.
bool buildBvh(const std::vector<BvhInputTri>& tris, const StkBvhOptions& opts, Bvh& out, std::string& err);

// Validates a skip-link array and derives packed nodes and triangles from it.
bool packBvh(const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni, const StkInstance* inst,
             uint32_t ninst, Bvh& bvh, std::string& err);

// Collapses the packed binary BVH into the compressed 8-wide layout and
// collects the emitter triangles.  Needs packBvh's output.
bool buildWide(const StkInstance* inst, uint32_t ninst, Bvh& bvh, std::string& err);

StkBvhOptions defaultBvhOptions();

} // namespace stk
